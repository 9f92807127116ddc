/* self_amd_io.h -- C ABI of libself_amd_io.so: reading HOPr mesh files (HDF5).
 *
 * Replaces the file access of Read_HOPr_Mesh2D (src/SELF_Mesh.f90:852-1016), which opens the
 * mesh file (Open_HDF5_serial, src/SELF_HDF5.f90:149-170) and reads the attributes nElems,
 * Ngeo, nBCs, nUniqueSides and the datasets ElemInfo, SideInfo, NodeCoords, GlobalNodeIDs
 * and BCType (ReadAttribute_HDF5_int32 :226-246, the serial ReadArray_HDF5 variants).  The
 * 3-D -> 2-D conversion that follows in the reference is selfamd_mesh2d_from_hopr (self_amd.h).
 *
 * A separate library so that libself_amd.so carries no HDF5 dependency.  Arrays are returned
 * in the file's row-major (C) order, e.g. SideInfo as (nSides, 5), NodeCoords as (nNodes, 3).
 * Also the model pickup files of WriteModel / ReadModel (Write_Model2D / Read_Model2D,
 * src/SELF_Model2D.f90:861-1099).
 *
 * Every entry point returns 0 on success, 1 for a bad argument and 5 for a file or HDF5
 * error; selfamd_io_last_error() describes the last failure.
 */
#ifndef SELF_AMD_IO_H
#define SELF_AMD_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *selfamd_io_last_error(void);

/* Integer attribute of the file's root group (ReadAttribute_HDF5_int32,
 * src/SELF_HDF5.f90:226-246). */
int selfamd_hopr_attribute_int(const char *path, const char *name, int32_t *value);

/* Rank (1 to 5) and dimensions (row-major, dims[0] slowest) of a dataset; dims holds 5. */
int selfamd_hopr_dataset_dims(const char *path, const char *name, int32_t *rank, int64_t *dims);

/* A whole dataset converted to int32 / double (the serial ReadArray_HDF5 variants,
 * e.g. ReadArray_HDF5_real_r2_serial, src/SELF_HDF5.f90:1928-1946); count is the number of
 * elements the caller's buffer holds and must equal the dataset's size. */
int selfamd_hopr_read_int(const char *path, const char *name, int32_t *out, int64_t count);
int selfamd_hopr_read_double(const char *path, const char *name, double *out, int64_t count);

/* WriteModel (Write_Model2D, src/SELF_Model2D.f90:861-1059, serial branch): host arrays in the
 * reference's Fortran layouts, written with their dimensions reversed as the reference's HDF5
 * writes show them: /quadrature/{xi, weights, dgmatrix, dmatrix, bmatrix},
 * /state/interior/solution (0:N,0:N,nVar,nEl), /state/boundary/solution (0:N,nVar,4,nEl),
 * /mesh/interior/x (2,0:N,0:N,1,nEl), /mesh/boundary/x (2,0:N,1,4,nEl), plus the root
 * attribute N, which Read_Model2D reads (:1078) and the reference never writes.  The
 * registers the reference also writes go in with selfamd_pickup_write_registers.  An existing
 * file is replaced. */
int selfamd_pickup_write(const char *path, int N, int nVar, int nEl, const double *xi, const double *weights,
                         const double *dgMatrix, const double *dMatrix, const double *bMatrix, const double *solution,
                         const double *solutionBoundary, const double *x, const double *xBoundary);

/* The remaining datasets of Write_Model2D (src/SELF_Model2D.f90:981-990), added to a file
 * that selfamd_pickup_write created: /state/interior/fluxDivergence (0:N,0:N,nVar,nEl),
 * /state/interior/flux (2,0:N,0:N,nVar,nEl), /state/boundary/flux (2,0:N,nVar,4,nEl),
 * /state/interior/solutionGradient (2,0:N,0:N,nVar,nEl) and
 * /state/boundary/solutionGradient (2,0:N,nVar,4,nEl), dimensions reversed as above.  A null
 * array is written as zeros (the reference's tendency path never fills the flux traces or the
 * solution gradient). */
int selfamd_pickup_write_registers(const char *path, int N, int nVar, int nEl, const double *fluxDivergence,
                                   const double *flux, const double *fluxBoundary, const double *solutionGradient,
                                   const double *solutionGradientBoundary);

/* ReadModel (Read_Model2D, src/SELF_Model2D.f90:1061-1099): checks the attribute N against N
 * and the solution's dimensions, then reads /state/interior/solution (0:N,0:N,nVar,nEl). */
int selfamd_pickup_read(const char *path, int N, int nVar, int nEl, double *solution);

#ifdef __cplusplus
}
#endif

#endif /* SELF_AMD_IO_H */
