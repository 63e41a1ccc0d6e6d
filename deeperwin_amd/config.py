"""Configuration records for the walker-parallel VMC inner loop.

Only the fields the MCMC step, the log|psi| forward and the local energy read are kept.  Defaults follow
the reference's pydantic defaults for the FermiNet-style model:

* MLP: tanh, uniform fan_avg weights, zero bias        (src/deeperwin/configuration.py:215-227)
* FermiNet embedding: 4 iterations, 256 / 32 widths    (configuration.py:277-375)
* FermiNet input features                              (configuration.py:634-652)
* envelope orbitals, full determinant, 32 dets         (configuration.py:688-699, 783-806)
* MCMC: Gaussian proposal (or the local-stepsize and
  Langevin proposals), target acceptance 0.5,
  stepsize clip [1e-2, 1.0], evaluation max_age 100    (configuration.py:952-1050)
* local energy: chunks of 64 walkers, forward_lap off  (configuration.py:1264-1266)
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import List, Optional, Union


@dataclass
class MLPConfig:
    activation: str = "tanh"
    init_bias_scale: float = 0.0
    init_weights_scale: str = "fan_avg"
    init_weights_distribution: str = "uniform"
    use_residual: bool = False
    use_layer_norm: bool = False


@dataclass
class EmbeddingConfigFermiNet:
    name: str = "ferminet"
    n_hidden_one_el: int = 256
    n_hidden_two_el: int = 32
    n_iterations: int = 4
    use_average_h_one: bool = True
    use_average_h_two: bool = True
    use_h_one: bool = True
    use_h_two_same_diff: bool = False
    use_el_ion_stream: bool = False
    use_schnet_features: bool = False


@dataclass
class EnvelopeOrbitalsConfig:
    envelope_type: str = "isotropic_exp"
    n_hidden: List[int] = field(default_factory=list)
    use_bias: bool = False
    initialization: str = "constant"


@dataclass
class OrbitalsConfigFermiNet:
    envelope_orbitals: EnvelopeOrbitalsConfig = field(default_factory=EnvelopeOrbitalsConfig)
    n_determinants: int = 32
    determinant_schema: str = "full_det"


@dataclass
class ModelConfigFermiNet:
    name: str = "ferminet"
    embedding: EmbeddingConfigFermiNet = field(default_factory=EmbeddingConfigFermiNet)
    orbitals: OrbitalsConfigFermiNet = field(default_factory=OrbitalsConfigFermiNet)
    mlp: MLPConfig = field(default_factory=MLPConfig)
    jastrow: Optional[object] = None
    use_el_el_cusp_correction: bool = False
    complex_wf: bool = False

    def validate(self) -> None:
        """Reject settings the native path does not implement (raises, as the reference does on
        unknown options, e.g. envelope_orbitals.py:81,85)."""
        e, o, m = self.embedding, self.orbitals, self.mlp
        if e.name != "ferminet":
            raise NotImplementedError(f"embedding {e.name!r} is not on the native path")
        if m.activation != "tanh" or m.use_layer_norm or m.use_residual:
            raise NotImplementedError("native MLP path is tanh, no layer norm, no global residual")
        if e.use_h_two_same_diff or e.use_el_ion_stream or e.use_schnet_features:
            raise NotImplementedError("only the default FermiNet symmetric features are implemented")
        if not (e.use_average_h_one and e.use_average_h_two and e.use_h_one):
            raise NotImplementedError("only the default FermiNet symmetric features are implemented")
        if o.determinant_schema != "full_det":
            raise NotImplementedError("only determinant_schema='full_det' is implemented")
        if o.envelope_orbitals.envelope_type != "isotropic_exp" or o.envelope_orbitals.n_hidden:
            raise NotImplementedError("only isotropic_exp envelopes without hidden backflow layers")
        if o.envelope_orbitals.use_bias:
            raise NotImplementedError("backflow output bias is off in the FermiNet model")
        if self.jastrow or self.use_el_el_cusp_correction or self.complex_wf:
            raise NotImplementedError("jastrow / cusp / complex wavefunctions are off in the FermiNet model")


@dataclass
class EmbeddingConfigDeepErwin4(EmbeddingConfigFermiNet):
    """The dpe4 embedding flags (configuration.py:422-441)."""
    name: str = "dpe4"
    use_el_ion_stream: bool = True
    use_h_two_same_diff: bool = True
    emb_dim: int = 32
    use_w_mapping: bool = True
    use_schnet_features: bool = True
    sum_schnet_features: bool = False
    use_average_h_two: bool = False
    use_linear_out: bool = False
    use_schnet_bias_feat: bool = True
    schnet_aggregation: str = "sum"
    use_deep_schnet_feat: bool = False
    h_one_correlation: int = 0


@dataclass
class InputFeatureConfigDPE4:
    """dpe4 input features (configuration.py:643-674): el-ion [dist, diff] concatenated into h_one, the
    el-el distance without differences, the ion embedding table ('lookup', 32 features)."""
    name: str = "dpe4"
    use_rbf_features: bool = False
    use_distance_features: bool = True
    use_el_ion_differences: bool = True
    use_el_el_differences: bool = False
    log_scale_distances: bool = False
    concatenate_el_ion_features: bool = True
    ion_embed_type: str = "lookup"
    n_ion_features: int = 32
    use_el_spin: bool = False
    coordinates: str = "cartesian"


@dataclass
class ModelConfigDeepErwin4(ModelConfigFermiNet):
    """The reference's default model (ModelConfigDeepErwin4, configuration.py:875-891).  Z_min / Z_max fix
    the ion embedding table (by default the smallest and largest charge of the molecule,
    model/wavefunction.py:310-326)."""
    name: str = "dpe4"
    embedding: EmbeddingConfigDeepErwin4 = field(default_factory=EmbeddingConfigDeepErwin4)
    features: InputFeatureConfigDPE4 = field(default_factory=InputFeatureConfigDPE4)
    Z_min: Optional[int] = None
    Z_max: Optional[int] = None

    def validate(self) -> None:
        e, f, o, m = self.embedding, self.features, self.orbitals, self.mlp
        if e.name != "dpe4" or f.name != "dpe4":
            raise NotImplementedError("ModelConfigDeepErwin4 runs the dpe4 embedding and features")
        if m.activation != "tanh" or m.use_layer_norm or m.use_residual:
            raise NotImplementedError("native MLP path is tanh, no layer norm, no global residual")
        flags = (e.use_el_ion_stream, e.use_h_two_same_diff, e.use_w_mapping, e.use_schnet_features,
                 not e.sum_schnet_features, not e.use_average_h_two, e.use_average_h_one, e.use_h_one,
                 not e.use_linear_out, e.use_schnet_bias_feat, e.schnet_aggregation == "sum",
                 not e.use_deep_schnet_feat, e.h_one_correlation == 0, e.emb_dim == 32, e.n_hidden_two_el == 32)
        if not all(flags):
            raise NotImplementedError("only the default dpe4 embedding flags are implemented")
        if (f.use_rbf_features or not f.use_distance_features or not f.use_el_ion_differences or
                f.use_el_el_differences or f.log_scale_distances or not f.concatenate_el_ion_features or
                f.ion_embed_type != "lookup" or f.n_ion_features != 32 or f.use_el_spin or f.coordinates != "cartesian"):
            raise NotImplementedError("only the default dpe4 input features are implemented")
        if o.determinant_schema != "full_det":
            raise NotImplementedError("only determinant_schema='full_det' is implemented")
        if o.envelope_orbitals.envelope_type != "isotropic_exp" or o.envelope_orbitals.n_hidden:
            raise NotImplementedError("only isotropic_exp envelopes without hidden backflow layers")
        if o.envelope_orbitals.use_bias:
            raise NotImplementedError("backflow output bias is off")
        if self.jastrow or self.use_el_el_cusp_correction or self.complex_wf:
            raise NotImplementedError("jastrow / cusp / complex wavefunctions are off in the dpe4 model")


@dataclass
class MCMCSimpleProposalConfig:
    """'normal', 'cauchy' or 'normal_one_el' (configuration.py:952-953)."""
    name: str = "normal"


@dataclass
class MCMCLangevinProposalConfig:
    """Local stepsize plus a drift towards the nuclei (configuration.py:956-963)."""
    name: str = "langevin"
    langevin_scale: float = 1.0
    r_min: float = 0.2
    r_max: float = 2.0


@dataclass
class LocalStepsizeProposalConfig:
    """Stepsize scaled by the distance to the closest nucleus, all electrons ('local') or one per step
    ('local_one_el') (configuration.py:966-975)."""
    name: str = "local"
    r_min: float = 0.1
    r_max: float = 1.0


@dataclass
class MCMCConfig:
    """Evaluation defaults (configuration.py:1045-1049); the configs fix the Gaussian init and
    stepsize 0.02 bohr instead of the reference's exponential init and 1e-2."""
    n_inter_steps: int = 20
    n_burn_in: int = 500
    max_age: int = 100
    stepsize_update_interval: int = 100
    n_walkers: int = 2048
    initialization: str = "gaussian"
    spin_initialization: str = "el_ion_mapping"
    target_acceptance_rate: float = 0.5
    min_stepsize_scale: float = 1e-2
    max_stepsize_scale: float = 1.0
    proposal: Union[MCMCSimpleProposalConfig, LocalStepsizeProposalConfig, MCMCLangevinProposalConfig] = field(
        default_factory=MCMCSimpleProposalConfig)
    p_spin_swap: float = 0.0
    p_spin_flip: float = 0.0
    stepsize: float = 0.02

    def validate(self) -> None:
        names = ("normal", "cauchy", "normal_one_el", "local", "local_one_el", "langevin")
        if self.proposal.name not in names:
            raise NotImplementedError(f"proposal {self.proposal.name!r} is not a proposal of the reference "
                                      f"sampler ({', '.join(names)})")
        if self.p_spin_swap > 0 or self.p_spin_flip > 0:
            raise NotImplementedError("spin swap / spin flip moves are not on the native path")


@dataclass
class ClippingConfig:
    """Local-energy clipping of the loss (configuration.py:1052-1058)."""
    name: str = "tanh"
    width_metric: str = "std"
    center: str = "mean"
    from_previous_step: bool = True
    clip_by: float = 5.0
    clip_imag_around_0: bool = False

    def validate(self) -> None:
        if self.name not in ("hard", "tanh"):
            raise ValueError(f"Unsupported config-value for optimization.clipping.name: {self.name}")
        if self.width_metric not in ("std", "mae"):
            raise NotImplementedError(f"Unknown clipping metric: {self.width_metric}")
        if self.center not in ("mean", "median"):
            raise NotImplementedError(f"Unknown clipping center: {self.center}")


@dataclass
class PhysicalConfig:
    """Molecule description (configuration.py:1537-1745): R in bohr, integer charges Z, electron counts."""
    R: List[List[float]]
    Z: List[int]
    n_electrons: Optional[int] = None
    n_up: Optional[int] = None
    el_ion_mapping: Optional[List[int]] = None
    name: Optional[str] = None
    E_ref: Optional[float] = None

    def __post_init__(self):
        from .physics import generate_el_ion_mapping

        if self.n_electrons is None:
            self.n_electrons = int(sum(self.Z))
        if self.n_up is None:
            # configuration.py:1740 with spin 0: extra electron goes to spin up
            self.n_up = (self.n_electrons + 1) // 2
        if self.el_ion_mapping is None:
            self.el_ion_mapping = generate_el_ion_mapping(self.R, self.Z, self.n_electrons, self.n_up)

    @property
    def n_dn(self) -> int:
        return self.n_electrons - self.n_up

    @property
    def n_ions(self) -> int:
        return len(self.Z)

    def get_basic_params(self):
        import numpy as np

        return self.n_electrons, self.n_up, np.array(self.R, dtype=np.float64), np.array(self.Z)


def replace(obj, **kw):
    return dataclasses.replace(obj, **kw)
