/*
 * hsgen_io - Fleur dump reader and H/S writer (host only, HDF5).
 *
 * Separate library (libhsgen_io.so, linked against libhsgen.so and HDF5) so
 * that the compute library carries no HDF5 dependency.  Layout of the files
 * is the one the reference reads and writes:
 *   input   fleur_dump.h5 written by Fleur's binmat.f90:63-474, read by
 *           FleurData (FleurData.cpp:175-458): /atoms {type, position[3],
 *           rotation[3][3] int}, /mesh {num_points, dx, rmt}, /lmax,
 *           /lmax_nonspherical, /energy_params (types x 9), /potential
 *           (types x max_mesh_pts, r V(r)), /real_basis, /rec_basis,
 *           /k_meshes {num_gpts, k_point[3], g_mesh[max_gpts][3]},
 *           /tmat/<type> (4 x n x n complex, lower triangles of t_uu, t_dd,
 *           t_ud, t_du), optionally Fleur's packed lower H and S in
 *           /matrix_hamilton/<k>, /matrix_overlap/<k> (binmat.f90:194-219);
 *   output  /matrix_hamilton/<k>, /matrix_overlap/<k> as n x n complex
 *           (HDF5Store.cpp:6-47), k 1-based.
 * Files written by hsgen_dump_write additionally carry the final per-atom
 * data the layout cannot express (group /hsgen: match values, rotation
 * matrices and final T per atom); the reader prefers it when present.
 *
 * All arrays are host memory; matrices are column-major complex128 as
 * interleaved doubles.  Functions return hsgen_status; the message of the
 * last failure on the calling thread is hsgen_io_last_error().
 */
#ifndef HSGEN_IO_H
#define HSGEN_IO_H

#include "hsgen.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hsgen_dump hsgen_dump;
typedef struct hsgen_store hsgen_store;

typedef struct {
    int32_t num_atoms, num_types;
    int32_t max_lmax;          /* max over types of /lmax */
    int32_t max_mesh_pts;
    int32_t num_kpts, max_gpts;
    double real_basis[9];      /* column-major, as FleurData holds it (FleurData.cpp:580-595) */
    double rec_basis[9];
    double omega;              /* det(real_basis) (FleurData.cpp:46) */
    int32_t num_hs;            /* k-points with Fleur H and S in the file */
    int32_t has_final;         /* 1 if /hsgen (final per-atom data) is present */
} hsgen_dump_info;

const char* hsgen_io_last_error(void);   /* never NULL */

/* Open and read a dump.  Replaces the FleurData constructor
 * (FleurData.cpp:175-212). */
hsgen_status hsgen_dump_open(const char* path, hsgen_dump** d);
hsgen_status hsgen_dump_close(hsgen_dump* d);
hsgen_status hsgen_dump_get_info(const hsgen_dump* d, hsgen_dump_info* info);

/* Per-type data (type 0-based): cutoffs, mesh, energy parameters
 * E[lmax+1], potential vr[mesh_n] and the raw /tmat blocks t4[4 n n complex]
 * (n = (lmax+1)^2).  Any output pointer may be NULL. */
hsgen_status hsgen_dump_get_type(const hsgen_dump* d, int32_t type, int32_t* lmax, int32_t* lnonsph,
                                 int32_t* mesh_n, double* dx, double* rmt, double* E, double* vr,
                                 double* t4);

/* Per-atom data: type (0-based), position (fractional), rotation as read
 * (9 doubles per atom, column-major view of the file's 3x3 integers). */
hsgen_status hsgen_dump_get_atoms(const hsgen_dump* d, int32_t* type, double* pos, double* rot);

/* k-point k (0-based): reduced k, G count and (if g != NULL) ng x 3 integer
 * G-vectors in file order.  Replaces get_kpoint/get_g_mesh
 * (FleurData.cpp:141-155). */
hsgen_status hsgen_dump_get_kpoint(const hsgen_dump* d, int32_t k, double* kpt, int32_t* ng, int32_t* g);

/* The system description for hsgen_system_create: radial solutions for
 * every (atom, l <= lmax) and final T-matrices (hsgen_radial_solve,
 * hsgen_tmat_assemble), or the /hsgen data when present.  Replaces
 * FleurData::populate_atomdata + AtomData::generate_matchdata +
 * augment_tmats (FleurData.cpp:38-68, AtomData.cpp:72-145).  The arrays
 * *desc points to are owned by d and valid until hsgen_dump_close. */
hsgen_status hsgen_dump_get_system(hsgen_dump* d, hsgen_system* desc);

/* Fleur's H and S of k-point k (0-based; datasets "<k+1>") as n*(n+1)/2
 * complex, packed lower triangle row by row.  *n receives the dimension;
 * H, S may be NULL to query it. */
hsgen_status hsgen_dump_get_hs_packed(const hsgen_dump* d, int32_t k, int64_t* n, double* H, double* S);

/* Unpack a row-wise packed lower triangle into the lower triangle of a
 * column-major n x n matrix with leading dimension ld (flapw_verify.py:68-82). */
hsgen_status hsgen_unpack_lower(int64_t n, const double* packed, double* M, int64_t ld);

/* Write a system and its k-points in the dump layout plus the /hsgen
 * extension (types[natoms] 0-based, amat column-major).  T of each type in
 * /tmat is the first atom's final T with the diagonal terms the reference's
 * loader adds removed (energy parameters are written as 0). */
hsgen_status hsgen_dump_write(const char* path, const hsgen_system* sys, const int32_t* types,
                              const double* amat, int32_t nk, const hsgen_kpoint* kps);

/* Add H and S of k-point k (0-based; datasets "<k+1>") to an existing dump
 * in Fleur's layout (binmat.f90:194-219): the lower triangles of the n x n
 * column-major matrices (leading dimension ld) packed row by row, as
 * hsgen_dump_get_hs_packed reads them.  Gives a synthetic dump the reference
 * H and S a comparison needs; an existing k-point is an argument error. */
hsgen_status hsgen_dump_write_hs_packed(const char* path, int32_t k, int32_t n, const double* H, const double* S,
                                        int64_t ld);
/* H/S output file.  Replaces HDF5Store (HDF5Store.cpp:6-47). */
hsgen_status hsgen_store_open(const char* path, hsgen_store** st);
hsgen_status hsgen_store_add_hs(hsgen_store* st, int32_t k, int32_t n, const double* H,
                                const double* S, int64_t ld);
hsgen_status hsgen_store_close(hsgen_store* st);

#ifdef __cplusplus
}
#endif

#endif /* HSGEN_IO_H */
