"""sph2grid -- MI355X (gfx950) implementation of SPHtoGrid.jl's particle
deposition, with the reference's host-side surface.

Public names mirror the Julia package (src/SPHtoGrid.jl:100-158 exports):
``mappingParameters``, ``sphMapping``, ``healpix_map``,
``filter_particles_in_image``, the ``part_weight_*`` presets and the SPHKernels
types.  All deposits run as HIP kernels through libsph2grid's C ABI
(include/sph2grid.h); there is no CPU fallback.
"""
from ._lib import Context, LibraryError, default_context, version  # noqa: F401
from .batch import BatchPlan, deposit_batch_device, map_batch  # noqa: F401
from .density import mass_density  # noqa: F401
from .distributed import (AllskySubfile, Subfile, distributed_allsky_map,  # noqa: F401
                          distributed_cic_map)
from .io import (get_map_grid_3D, read_allsky_fits_image, read_fits_image,  # noqa: F401
                 read_healpix_fits, read_vtk_image, save_healpix_fits, write_fits_image,
                 write_vtk_image)
from .healpix import BoundsError, filter_sort_particles, healpix_map, healpix_maps  # noqa: F401
from .kernels import (AbstractSPHKernel, Cubic, Quintic, WendlandC2,  # noqa: F401
                      WendlandC4, WendlandC6, WendlandC8)
from .mapping import (center_particles, deposit, filter_particles_in_image,  # noqa: F401
                      sphMapping)
from .parameters import ErrorException, MappingParameters, mappingParameters  # noqa: F401
from .tsc import tsc_mapping  # noqa: F401
from .weights import (part_weight_emission, part_weight_one,  # noqa: F401
                      part_weight_physical, part_weight_spectroscopic,
                      part_weight_XrayBand)

__all__ = [
    "mappingParameters", "MappingParameters", "ErrorException", "sphMapping", "deposit",
    "center_particles", "filter_particles_in_image", "part_weight_one",
    "part_weight_physical", "part_weight_emission", "part_weight_spectroscopic",
    "part_weight_XrayBand", "AbstractSPHKernel", "Cubic", "Quintic", "WendlandC2",
    "WendlandC4", "WendlandC6", "WendlandC8", "Context", "LibraryError",
    "default_context", "version", "healpix_map", "healpix_maps", "filter_sort_particles",
    "BoundsError",
    "tsc_mapping", "map_batch", "deposit_batch_device", "BatchPlan", "distributed_cic_map",
    "distributed_allsky_map", "Subfile", "AllskySubfile", "write_fits_image",
    "read_fits_image", "read_allsky_fits_image", "save_healpix_fits", "read_healpix_fits",
    "write_vtk_image", "read_vtk_image", "get_map_grid_3D", "mass_density",
]
