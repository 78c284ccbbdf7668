"""MI355X-native tiled Gaussian-splatting frame pipeline.

The package name starts with a digit, so import it with
``importlib.import_module("3dgs_tile_webgpu_amd")``.

Public API (mirrors the reference's pass API for the frame path):
  gaussian_pass / GaussianPass   -- one camera's pass over the C ABI
  BatchedPass                    -- many cameras in one captured launch,
                                    with the launch-shape autotuner
  LearnedCull                    -- the MLP cull fused into projection,
                                    trained on the GPU from plain frames
  PackedScene, workload_scene    -- packed buffers and synthetic workloads
  read_ply, pack_scene           -- canonical 3DGS PLY files into the store's
                                    slot buffers (holes beyond the cloud)
  ActiveSlots, check_data        -- GaussianData's active slot list (staged
                                    removals/additions, committed ranges) and
                                    its buffer checks
  PerspectiveCamera              -- Three.js camera conventions
"""
from ._native import GsError, GsRangeError, GsTypeError, lib  # noqa: F401
from .active_slots import ActiveSlots, check_data  # noqa: F401
from .batch import BatchedPass  # noqa: F401
from .camera import PerspectiveCamera, elements, orbit_cameras  # noqa: F401
from .gaussian_pass import (  # noqa: F401
    GaussianPass,
    IntersectionCapacityPolicy,
    frame_for,
    gaussian_pass,
    object_frames_for,
    validate_options,
)
from .learned_cull import LearnedCull, psnr  # noqa: F401
from .scene import (BATCH_RESOLUTIONS, WORKLOADS, PackedScene, Workload, generate,  # noqa: F401
                    pack_scene, read_ply, spec_for, workload_scene)
