/*
 * mosaic_h5.h -- C ABI of the HDF5 container layer around the hot path: the
 * dxchange files the reference pipeline reads and writes between its steps
 * (proj.h5 / data.h5: /exchange/data (nproj, ny, nx) float32 as a virtual
 * dataset over bank files, /exchange/theta float32 degrees).
 *
 * Replaces the h5py code of iohdf5/dxchange_hdf5_chunks.py: tomo_info (:99),
 * tomo_initx (:223) with its banking plan (:182), the bank writes of
 * tomo_writex / write_vchunkx (:332, :443-520, :681) and the box reads of
 * read_projs_vchunkx / read_slices_vchunkx (:524, :595, :636), plus the theta
 * dataset of step2_radon.py:129-137.  Built on the HDF5 C library (1.10).
 *
 * Conventions: every call returns MM_H5_OK (0) or a negative code, with the
 * calling thread's message in mm_h5_last_error().  Files are opened and
 * closed inside each call.  Boxes are (start[3], count[3]) in the dataset's
 * (nproj, ny, nx) index space; host buffers are C-contiguous count[0] x
 * count[1] x count[2] float32.
 */
#ifndef MOSAIC_H5_H
#define MOSAIC_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define MM_H5_OK 0
#define MM_H5_ERR_ARG (-1) /* bad argument or box outside the dataset */
#define MM_H5_ERR_IO (-2)  /* HDF5 reported an error (message has the call) */

#define MM_H5_STYPE_PROJ 0  /* bank files split along axis 0 (projections) */
#define MM_H5_STYPE_SLICE 1 /* bank files split along axis 1 (slices) */

typedef struct mm_h5_info {
  long long shape[3];        /* nproj, ny, nx */
  int is_virtual;            /* 1: virtual dataset over bank files */
  long long chunks[3];       /* HDF5 chunk shape (recorded one for a virtual dataset; 0: contiguous) */
  long long vchunks[3];      /* the writer's vchunk shape (attributes vchunks_0..2) */
  int nbanks_per_svchunk;    /* attribute nbanks_per_svchunk (1 if absent) */
  int stype;                 /* MM_H5_STYPE_* (attribute stype) */
} mm_h5_info;

const char* mm_h5_last_error(void);

/* tomo_info: shape, layout and the writer's attributes of /exchange/data. */
int mm_h5_info_get(const char* path, mm_h5_info* info);

/* Box read of /exchange/data into dst (virtual datasets read through to their bank files). */
int mm_h5_read(const char* path, const long long* start, const long long* count, float* dst);

/* tomo_initx: creates the bank files <dir>/<base>/<base>_data_%06d<ext> (chunked float32
 * /exchange/data, no fill) and the master file (libver latest) whose /exchange/data is a virtual
 * dataset over them with fill value -5 and the attributes nbanks_per_svchunk, stype, vchunks_*,
 * chunks_*.  chunks may be NULL (default: (1, vchunks[1], vchunks[2]) for proj,
 * (vchunks[0], 1, vchunks[2]) for slice).  Existing files are replaced. */
int mm_h5_create(const char* path, const long long* shape, const long long* vchunks, int stype, int nbanks,
                 const long long* chunks);

/* Writes a box of /exchange/data: into the bank files it covers for a virtual dataset, else the dataset. */
int mm_h5_write(const char* path, const long long* start, const long long* count, const float* src);

/* /exchange/theta (float32, degrees): replaced on write; mm_h5_read_theta returns its length in *n
 * (theta may be NULL to query it; otherwise *n is the capacity of theta). */
int mm_h5_write_theta(const char* path, const float* theta_deg, long long n);
int mm_h5_read_theta(const char* path, float* theta_deg, long long* n);

#ifdef __cplusplus
}
#endif

#endif /* MOSAIC_H5_H */
