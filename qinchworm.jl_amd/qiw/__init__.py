"""qiw -- MI355X (gfx950) quasi Monte Carlo diagram evaluation for the inchworm solver.

Host-side Python mirror of QInchworm.jl's inchworm!/correlator_2p API over the C ABI of
libqiw.so (include/qiw.h).  Setup (model construction, topologies, Sobol direction numbers)
happens here; all sampling and diagram evaluation runs in the native library on the GPU.
"""
from . import abi, diagrammatics, model, randomization, sobol  # noqa: F401
from .randomization import RandomizationParams, RandomSeq, RequestStdDev  # noqa: F401
from .inchworm import (autotune, context, correlator_2p, density_matrix, diff_inchworm, get_tuning, inchworm,  # noqa: F401
                       inchworm_step, inchworm_step_bare, normalize, partition_function, scramble_dims,
                       set_graphs, set_points, set_scrambles, set_tuning, stats)
from .warp import (group_index, warp_apply, warp_enable, warp_init, warp_loss_grad,  # noqa: F401
                   warp_params, warp_train)
