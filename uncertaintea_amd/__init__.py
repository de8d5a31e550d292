"""uncertaintea_amd: the MI355X (gfx950) device backend of UncertainTea's batched
density and batched MCMC path, behind the C ABI of libuncertaintea_hip."""
from . import diagnostics, synthetic
from ._lib import DeviceError, NonFiniteInitError, device_available, philox4x32, philox_fill
from .device import (DeviceBatchedWorkspace, device_batched_logjoint, device_batched_logjoint_gradient,
                     device_batched_logjoint_gradient_inplace, device_batched_logjoint_inplace,
                     device_lowering_report)
from .export import (parameter_names, pointwise_loglikelihood, posterior_array, to_arviz_dict,
                     write_crossppl_run)
from .flow import AffineCouplingFlow, FullRankGaussian
from .inference import HMCChains, ProgressInfo, batched_chees, batched_hmc, batched_meads, batched_nuts
from .plan import Expr, PlanBuilder, PlanModel, VectorLatent
from .models import (GaussianProcessPoisson, HierarchicalNormalNC, LogisticRegression, LogisticRegressionBF16,
                     NealsFunnel, TeaModel)
from .vi import ADVIResult, batched_advi, variational_covariance, variational_mean, variational_samples

__all__ = [
    "Expr", "PlanBuilder", "PlanModel", "VectorLatent",
    "ADVIResult", "AffineCouplingFlow", "FullRankGaussian", "batched_advi", "variational_covariance", "variational_mean",
    "variational_samples", "parameter_names", "pointwise_loglikelihood", "posterior_array", "to_arviz_dict", "write_crossppl_run",
    "DeviceBatchedWorkspace", "GaussianProcessPoisson", "DeviceError", "HMCChains", "HierarchicalNormalNC", "LogisticRegression",
    "LogisticRegressionBF16", "NealsFunnel", "NonFiniteInitError", "TeaModel", "batched_chees", "batched_hmc", "batched_meads", "batched_nuts", "device_available",
    "device_batched_logjoint", "device_batched_logjoint_gradient", "device_batched_logjoint_gradient_inplace",
    "device_batched_logjoint_inplace", "device_lowering_report", "ProgressInfo", "diagnostics", "philox4x32", "philox_fill", "synthetic",
]
