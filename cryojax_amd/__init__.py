"""MI355X-native (gfx950 HIP) Fourier-slice image formation and pose search.

A drop-in for cryoJAX's Fourier-voxel image-formation hot path:
`cryojax_amd.simulator` mirrors `cryojax.simulator` (make_image_model,
simulate(), FourierVoxelGridVolume, EulerAnglePose, QuaternionPose,
AstigmaticCTF, ContrastTransferTheory, BasicImageConfig,
FourierSliceExtraction) and every image is formed by the HIP kernels behind
the C ABI in include/cryomi.h.
"""

from . import _lib, rotations, synthetic  # noqa: F401

__version__ = "0.1.0"
