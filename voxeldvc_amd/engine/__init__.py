"""Drop-in equivalents of voxelDVC's engine modules (geometry_dvc,
kernels_dvc, correlate_gpu) backed by the gfx950 library libvdvc.so."""
