"""MI355X implementation of baby_shark's implicit-modeling path (src/voxel).

Public names mirror ``baby_shark::voxel::prelude``; see ``voxel.py``.
"""
from ._lib import BSError, load as load_library  # noqa: F401
from .voxel import Context, DualContouringMesher, MarchingCubesMesher, MeshToVolume, Volume, VolumeBuilder, VolumeExport, batch_mesh_to_volume, default_context, read_stl, release_cached_memory, volume_from_voxels, write_stl, MeshingMethod, VoxelRemesher, TileClassifier, tile_samples  # noqa: F401
from . import meshgen  # noqa: F401
