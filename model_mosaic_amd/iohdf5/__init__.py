"""HDF5 containers (dxchange layout) of the reference pipeline's files, on the HDF5 C library."""
