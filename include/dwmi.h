/*
 * dwmi.h — C ABI of the MI355X-native (gfx950) walker-parallel VMC inner loop.
 *
 * The library evaluates the FermiNet-style wavefunction of the reference (mdsunivie/deeperwin), its
 * Metropolis-Hastings chain and its local energy on one GPU.  Every entry point below replaces one
 * Python call of the reference; the reference interface is cited (file:line under src/deeperwin/).
 *
 * Conventions
 *  - All array arguments are DEVICE pointers owned by the caller, row-major, float32 unless stated.
 *    r [B, n_el, 3]; R [n_ion, 3]; Z [n_ion] (charges as float).
 *  - Calls are ordered on the given HIP stream and never synchronise the host.
 *  - A model handle may be used from one stream at a time; different handles are independent.
 *  - Errors return a non-zero dw_status; dw_last_error() gives a thread-local message.  Non-finite
 *    numbers are values, not errors: they are reported per walker through out_flags.
 *  - The handle owns its parameter copy and a workspace that grows on demand (hipMalloc).
 */
#ifndef DWMI_H
#define DWMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int dw_status;
#define DW_OK 0
#define DW_ERR_INVALID 1
#define DW_ERR_HIP 2
#define DW_ERR_STATE 3

typedef struct dw_model dw_model;
typedef void* dw_stream; /* a hipStream_t; NULL is the default stream */

/* Precision forms.  DW_FP32: every GEMM in FP32 on MFMA, determinant stage in FP64.  DW_BF16X3:
 * embedding and orbital GEMMs on BF16 MFMA with FP32 accumulation, operands split into hi+lo BF16 pairs
 * (three products).  DW_BF16: one BF16 product.  DW_FP64: the whole pass in FP64 (the reference's
 * float_precision="float64" mode, configuration.py:1893).  With fp64_cond_threshold > 0, walkers whose
 * conditioning estimate exceeds it are recomputed by dw_local_energy in FP64 and flagged DW_FLAG_FP64_DET. */
#define DW_FP32 0
#define DW_BF16X3 1
#define DW_BF16 2
#define DW_FP64 3

/* Per-walker flag bits written to out_flags. */
#define DW_FLAG_NONFINITE 1u     /* log|psi|^2 or E_loc is NaN/inf                          */
#define DW_FLAG_FP64_DET 2u      /* conditioning above threshold: recomputed in FP64         */
#define DW_FLAG_FORCED_ACCEPT 4u /* MCMC: accepted because walker age reached max_age       */

typedef struct {
  int32_t n_el;    /* electrons                                  */
  int32_t n_up;    /* spin-up electrons (the first n_up rows)    */
  int32_t n_ion;   /* nuclei                                     */
  int32_t n_det;   /* full determinants                          */
  int32_t width1;  /* 1-el stream width (256)                    */
  int32_t width2;  /* 2-el stream width (32)                     */
  int32_t n_iter;  /* FermiNet iterations (4)                    */
  int32_t precision;       /* DW_FP32 / DW_BF16X3 / DW_BF16 / DW_FP64 */
  float fp64_cond_threshold; /* walker conditioning above which dw_local_energy recomputes in FP64 (<=0: never) */
  int32_t embedding;       /* DW_EMBEDDING_FERMINET (0) or DW_EMBEDDING_DPE4 (1)                      */
  int32_t z_min;           /* dpe4: nuclear charge of the ion embedding table's first row (Z_min)      */
  int32_t n_ion_types;     /* dpe4: rows of the ion embedding table (Z_max - Z_min + 1)                */
} dw_model_desc;

/* Embeddings.  FERMINET: the FermiNet-style model (ferminet_embedding.py:207-267 with the FermiNet flags).
 * DPE4: the reference's default model (ModelConfigDeepErwin4, configuration.py:875-891, 422-441, 666-674):
 * same/diff-spin pair streams started from the el-el distance, an el-ion stream, SchNet-style convolutions
 * e_ee[i] = sum_j w_same|diff(pair[i, j]) * tanh(h W_map + b)[j] and e_ei[i] = sum_I el_ion[i, I] *
 * tanh(h_ion W + b)[I] with h_ion = embeddings[Z_I - Z_min] (ferminet_embedding.py:52-188,
 * input_features.py:176-183); its leaves are listed at dw_model_set_param.  width2 (= emb_dim) must be 32.
 * dw_param_grad covers both embeddings (for DPE4 also the convolution maps, the pair and el-ion streams, the
 * ion maps and the rows of the ion table that the molecule's nuclei use; unused rows get zero). */
#define DW_EMBEDDING_FERMINET 0
#define DW_EMBEDDING_DPE4 1

/* Create a model for one (n_el, n_up, n_ion) shape.  Replaces build_log_psi_squared's model
 * construction (model/wavefunction.py:262-299). */
dw_status dw_model_create(const dw_model_desc* desc, dw_model** out);
dw_status dw_model_destroy(dw_model* m);

/* Upload one parameter leaf from HOST memory (float32, row-major).  `module_path` is the haiku module
 * path and `leaf` the leaf name (w, b, alpha_up, alpha_dn, weights_up, weights_dn, embeddings); paths are
 * matched by suffix, so 'wf/ferminet_embedding/h_el_0/linear_0' and '.../~/h_el_0/linear_0' are both
 * accepted.  dpe4 leaves: symm_features_{l}/convolutional_features/{w_same,w_diff,h_map,h_ion_map}/linear_0,
 * h_same_{l}/linear_0, h_diff_{l}/linear_0, h_el_ion_{l}/linear_0 (w, b) and input/h_ion (embeddings).
 * Replaces handing `params` (a haiku dict) to log_psi_sqr (model/wavefunction.py:293; leaves at
 * mlp.py:57-59, envelope_orbitals.py:101-104). */
dw_status dw_model_set_param(dw_model* m, const char* module_path, const char* leaf, const float* host_data,
                             const int64_t* shape, int32_t ndim);
/* Returns 1 when every leaf has been set. */
int32_t dw_model_ready(const dw_model* m);

/* log|psi|^2 and phase per walker: log_psi_sqr(params, n_up, n_dn, r, R, Z, fixed_params)
 * -> (phase, log_psi_sqr)  (model/wavefunction.py:118-134, 293). */
dw_status dw_logpsi(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                    float* out_log_psi_sqr, float* out_phase, dw_stream stream);

/* Local energy per walker (kinetic via a forward-mode Laplacian over 3*n_el directions, plus Coulomb
 * potential): get_local_energy(params, spin_state, r, R, Z, fixed_params) -> E_loc[B]
 * (hamiltonian.py:272-291, 198-269, 17-39).  out_log_psi_sqr may be NULL; out_flags may be NULL. */
dw_status dw_local_energy(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                          float* out_eloc, float* out_log_psi_sqr, uint32_t* out_flags, dw_stream stream);

/* Chain scalars kept on the device so a run of steps needs no host synchronisation. */
typedef struct {
  float stepsize;
  float acc_rate;
  int32_t step_nr;
  int32_t _pad;
} dw_chain_scalars;

/* Proposal rules of the sampler (the `proposal` config of MCMCConfig, configuration.py:952-975, 1017-1021). */
#define DW_PROPOSAL_NORMAL 0   /* r' = r + stepsize xi                                  (mcmc.py:175-180) */
#define DW_PROPOSAL_LOCAL 1    /* per-electron stepsize from the closest nucleus        (mcmc.py:218-236) */
#define DW_PROPOSAL_LANGEVIN 2 /* local stepsize plus a drift towards the nuclei        (mcmc.py:255-284) */
#define DW_PROPOSAL_NORMAL_ONE_EL 3 /* only electron step_nr mod n_el moves, by stepsize xi (mcmc.py:183-192) */
#define DW_PROPOSAL_LOCAL_ONE_EL 4  /* that electron with its local stepsize              (mcmc.py:239-252) */

/* Select the proposal of dw_mcmc_steps / dw_mcmc_steps_noalloc / dw_mcmc_propose for this model (default
 * DW_PROPOSAL_NORMAL).  For LOCAL and LANGEVIN each electron moves with s = stepsize * clip(d, r_min, r_max),
 * d its distance to the closest nucleus; LANGEVIN adds g s^2 with g = -langevin_scale sum_I Z_I (r - R_I) /
 * |r - R_I|.  The accept test becomes exp(lp' - lp + log q) > u with log q the log ratio of the reverse to
 * the forward proposal density.  The ONE_EL forms move only electron k = step_nr mod n_el of every walker
 * (step_nr of the chain record before the step) and take 3 draws per walker and step: normals
 * [n_steps, B, 3]; LOCAL_ONE_EL's log q is that electron's term alone, as in the reference.  The Cauchy
 * proposal (mcmc.py:195-200) is DW_PROPOSAL_NORMAL fed Cauchy draws.  The reference's defaults: local r_min
 * 0.1, r_max 1; langevin r_min 0.2, r_max 2, langevin_scale 1 (configuration.py:956-975).  Needs
 * 0 < r_min <= r_max.  Replaces the choice of self.propose in MetropolisHastingsMonteCarlo.__init__
 * (mcmc.py:330-342). */
dw_status dw_model_set_proposal(dw_model* m, int32_t kind, float r_min, float r_max, float langevin_scale);

/* One proposal of the selected rule without evaluating the wavefunction: out_r [B, n_el, 3] and
 * out_log_q [B] (0 for the symmetric rules) from r [B, n_el, 3], normals [B, n_el, 3] ([B, 3] for the ONE_EL
 * rules) and the chain's stepsize (and step_nr).  Replaces the propose functions `_propose_normal`,
 * `_propose_normal_one_el`, `_propose_step_local_stepsize`, `_propose_step_local_stepsize_one_el` and
 * `_propose_step_local_approximated_langevin` (mcmc.py:175-192, 218-284) with supplied draws. */
dw_status dw_mcmc_propose(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                          const float* normals, const dw_chain_scalars* chain, float* out_r, float* out_log_q,
                          dw_stream stream);

/* n_steps Metropolis-Hastings steps with the model's proposal rule (dw_model_set_proposal) and SUPPLIED
 * draws: normals [n_steps, B, n_el, 3] ([n_steps, B, 3] for the ONE_EL rules), uniforms [n_steps, B].  r, log_psi_sqr and walker_age are updated
 * in place; chain (device dw_chain_scalars) carries stepsize / acc_rate / step_nr; out_acc [n_steps]
 * receives each step's acceptance fraction (may be NULL).  log_psi_sqr is re-evaluated on entry.
 * Replaces MetropolisHastingsMonteCarlo.run_mcmc_steps (mcmc.py:325-327, 345-406). */
dw_status dw_mcmc_steps(dw_model* m, float* r, int32_t n_walkers, const float* R, const float* Z,
                        float* log_psi_sqr, int32_t* walker_age, const float* normals, const float* uniforms,
                        int32_t n_steps, dw_chain_scalars* chain, int32_t max_age,
                        int32_t stepsize_update_interval, float target_acceptance, float min_stepsize,
                        float max_stepsize, float* out_acc, uint32_t* out_flags, dw_stream stream);

/* dw_mcmc_steps without any device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as dw_mcmc_steps, but every device buffer of the call
 * (activations of the value pass, proposals, acceptance counters) is carved from one static device array
 * of DW_NOALLOC_SCRATCH_BYTES per GPU, held from library load to process end; the call makes no hipMalloc
 * or hipFree.  Fails with DW_ERR_INVALID when n_walkers do not fit that array (the message states how
 * many fit) and with DW_ERR_STATE when the weights are not on the device yet (call dw_logpsi, dw_mcmc_steps
 * or dw_local_energy once after the last dw_model_set_param).  Every call on one GPU shares the array:
 * calls that overlap in time on one GPU race.  It does not update the dw_model_conditioning record. */
#define DW_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
dw_status dw_mcmc_steps_noalloc(dw_model* m, float* r, int32_t n_walkers, const float* R, const float* Z,
                                float* log_psi_sqr, int32_t* walker_age, const float* normals, const float* uniforms,
                                int32_t n_steps, dw_chain_scalars* chain, int32_t max_age,
                                int32_t stepsize_update_interval, float target_acceptance, float min_stepsize,
                                float max_stepsize, float* out_acc, uint32_t* out_flags, dw_stream stream);

/* ---- Parameter gradient (the optimisation step's tangent) ----
 * G = sum_b c_b d log_psi_sqr(r_b) / d theta for every parameter leaf, with the caller's per-walker
 * weights c (device, n_walkers floats).  With c_b = (E_clip_b - mean E_clip) / B this is the gradient that
 * the reference's total-energy JVP produces (optimization/loss_function.py:115-156: jnp.dot of the
 * log_psi_sqr tangents with the clipped deviations, divided by the batch size).  The result is written as
 * float32 to out_grad (device, dw_param_count floats): the leaves one after another in the order of
 * dw_param_layout, each row-major in the reference's shape (w: [in, out], b: [out], envelope [n_ion, D n]).
 * out_log_psi_sqr (may be NULL) receives log_psi_sqr of the walkers.  Reverse mode through the value pass:
 * determinants inverted in FP64, every reduction in a fixed order.  Uses its own device workspace (not the
 * caller-owned one of dw_model_set_workspace).  NaN walker weights give NaN gradients, as in the reference.
 * Up to 52 electrons. */
dw_status dw_param_grad(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                        const float* walker_weights, float* out_grad, float* out_log_psi_sqr, dw_stream stream);
/* Clipped local energies of the loss (optimization/loss_function.py:12-110, one device): clipping center
 * nanmean (center_median = 0) or nanmedian (1, at most 8192 walkers), width sqrt(nanmean (E - c)^2) (std,
 * width_mae = 0) or nanmean |E - c| (mae, 1) times clip_by, clip tanh (clip_tanh = 1: c + tanh((E - c) / w)
 * w) or hard (0: clip to [c - w, c + w]).  With from_previous_step and state_in (device, 2 doubles: center,
 * width) the previous step's state is used, else it is computed from E.  Writes the per-walker weights
 * c_b = (E_clip_b - mean E_clip) / B for dw_param_grad, optionally E_clip, and out_stats (device, 8 doubles):
 * [loss = nanmean E_clip, E_mean, E_var, E_var_clipped, next center, next width, center used, width used];
 * the next state is computed from E_clip.  FP64, fixed summation order. */
dw_status dw_clip_energies(const float* e_loc, int32_t n_walkers, int32_t clip_tanh, int32_t center_median,
                           int32_t width_mae, float clip_by, int32_t from_previous_step, const double* state_in,
                           float* out_weights, float* out_e_clipped, double* out_stats, dw_stream stream);
/* Number of floats of the packed parameter gradient. */
int64_t dw_param_count(const dw_model* m);
/* Leaf `index` (0 .. number of leaves - 1) of the packed gradient: key "<module suffix>:<leaf>" (e.g.
 * "h_el_0/linear_0:w", "envelope_orbitals:alpha_up") into key[key_len], its float offset and its shape
 * (shape[1] = 1 for vectors).  DW_ERR_INVALID past the last leaf. */
dw_status dw_param_layout(const dw_model* m, int32_t index, char* key, int32_t key_len, int64_t* offset,
                          int64_t* shape);

/* ---- JAX random streams (jax 0.4.23's default threefry2x32 PRNG; keys are 2 uint32 words on the device) ----
 * so that the sampler replays the reference's draws (mcmc.py:43, 90, 175-180, 345-378).
 * dw_threefry_split: jax.random.split(key, num) -> out [num, 2].
 * dw_threefry_normal: jax.random.normal(key, shape) for `count` = prod(shape) float32 values (row-major).
 * dw_threefry_mcmc_draws: the sampler's per-walker stream for n_steps steps: at each step the walker key k
 *   gives s = split(k)[1] and the next key split(k)[0]; normals[step, b] = normal(s, (n_el, 3)) (the
 *   proposal, mcmc.py:178) and uniforms[step, b] = uniform(s, ()) (the accept threshold, mcmc.py:360-361,
 *   the same subkey).  keys_in [B, 2] -> keys_out [B, 2] (must not alias).  The normals' erfinv is the
 *   single-precision polynomial of M. Giles (GPU Computing Gems 2011), as XLA's; the uniforms are exact. */
dw_status dw_threefry_split(const uint32_t* key, int32_t num, uint32_t* out, dw_stream stream);
dw_status dw_threefry_normal(const uint32_t* key, int64_t count, float* out, dw_stream stream);
dw_status dw_threefry_mcmc_draws(const uint32_t* keys_in, uint32_t* keys_out, int32_t n_walkers, int32_t n_el,
                                 int32_t n_steps, float* normals, float* uniforms, dw_stream stream);

/* E_mean and E_var of a batch of local energies in FP64: out_mean_var (device, 2 doubles) receives
 * nanmean(E_loc) and nanmean(|E_loc - E_mean|^2).  NaN walkers are dropped, +-inf kept, as jnp.nanmean
 * does.  Replaces get_observables' energy metrics (optimization/evaluation.py:143-151; pmean over one
 * device is the identity).  The summation order is fixed, so the result does not vary between runs. */
dw_status dw_energy_moments(const float* e_loc, int32_t n_walkers, double* out_mean_var, dw_stream stream);

/* Copy the per-walker conditioning estimate of the last dw_logpsi or dw_local_energy call into out (device,
 * n_walkers floats, at most that call's walkers, ordered on `stream`): max_d |A_d|_1 |A_d^-1|_1 (local
 * energy; pivot ratio in dw_logpsi) times sum_d |s_d e^{l_d-m}| / |sum_d s_d e^{l_d-m}|, for every chunk of
 * a chunked dw_local_energy.  Walkers recomputed in FP64 keep the FP32 estimate that flagged them. */
dw_status dw_model_conditioning(dw_model* m, float* out, int32_t n_walkers, dw_stream stream);

/* Bytes of device workspace the model holds after sizing for n_walkers in the given mode
 * (0 = log|psi| / MCMC, 1 = local energy). */
size_t dw_workspace_size(const dw_model* m, int32_t n_walkers, int32_t mode);

/* Caller-owned workspace (no counterpart in the reference; the analogue of an XLA FFI scratch buffer).
 * When set, the large activation buffers of every pass are carved from [device_ptr, device_ptr + bytes)
 * instead of being allocated by the library, so a caller that recycles buffers (a serving loop, a
 * graph) avoids device allocation and release per model.  bytes must be >= dw_workspace_size for every
 * (n_walkers, mode) the model is then called with, else the call fails; the pointer must be 256-byte
 * aligned and stay valid while the model may use it.  Calls on one model must not overlap (one stream).
 * device_ptr = NULL returns to library-owned workspace. */
dw_status dw_model_set_workspace(dw_model* m, void* device_ptr, size_t bytes);

/* Cap on walkers processed per internal pass of dw_local_energy (0 = no cap).  Bounds the workspace. */
dw_status dw_model_set_chunk(dw_model* m, int32_t max_walkers_per_pass);

/* Launch-shape choices of one model (no counterpart in the reference: XLA picks its own tiles).  Set by
 * the autotuner in deeperwin_amd/tuning.py per electron-count bucket; -1 / 0 keep the built-in heuristic.
 * FP32 GEMM tiles: 0 = 64x64, 1 = 64x128, 2 = 128x128 per 256-thread workgroup; fwd_main_tile also takes
 * 4 = 112 rows x 256 columns (N = 256; the heuristic's choice where it would take 64x64 and that gives at
 * least 192 row tiles). */
typedef struct dw_tuning {
  int32_t fwd_main_tile;  /* 1-el embedding GEMMs of the value pass (3: 96 x 128 tiles, 16x16 MFMA)   */
  int32_t fwd_orb_tile;   /* orbital GEMMs of the value pass                                         */
  int32_t fwd_mu_split;   /* K-split (1, 2 or 4) of the spin-mean GEMMs of the value pass (FP32 form) */
  int32_t lap_orb_tile;   /* orbital GEMMs of the Laplacian pass                                     */
  int32_t lap_mu_tile;    /* spin-mean GEMMs of the Laplacian pass                                   */
  int32_t lap_electrons;  /* electrons per workgroup of the fused Laplacian GEMM (E * Cp <= 192)      */
  int32_t lap_cols;       /* output columns per workgroup of that GEMM: 0 / 128, or 256 (FP32 form,
                             E * Cp <= 96; larger tiles fall back to 128)                             */
} dw_tuning;
dw_status dw_model_set_tuning(dw_model* m, const dw_tuning* t);
dw_status dw_model_get_tuning(const dw_model* m, dw_tuning* out);

/* ---- Measurement hooks (no counterpart in the reference; used by the benchmark harness) ----
 * When timing is enabled the library brackets selected launches with HIP events recorded on the
 * launch stream; dw_model_timer_query synchronises those events and returns the summed duration and the
 * number of bracketed launches of one region kind since timing was (re-)enabled. */
#define DW_TIMER_GEMM_1EL_LAP 0 /* 1-el embedding GEMM, Laplacian mode, iterations >= 1 (one launch each) */
#define DW_TIMER_DET_LAP 1      /* determinant kernel, Laplacian mode                                   */
#define DW_TIMER_FORWARD 2      /* one value-only wavefunction pass (all its kernels)                   */
#define DW_TIMER_LOCAL_ENERGY 3 /* one Laplacian pass (all its kernels)                                 */
#define DW_TIMER_KINDS 4
dw_status dw_model_set_timing(dw_model* m, int32_t enable);
dw_status dw_model_timer_query(dw_model* m, int32_t kind, double* total_ms, int32_t* launches);

/* Thread-local message of the last failing call on this thread. */
const char* dw_last_error(void);

/* Library version string. */
const char* dw_version(void);

#ifdef __cplusplus
}
#endif
#endif /* DWMI_H */
