// Synthetic linear systems with the structure of StiffGIPC's Newton-step
// Hessians: lumped mass / dt^2 plus Stable Neo-Hookean element Hessians made
// PSD by eigenvalue clamping (femEnergy.cu:1641-1717 builds the same
// projection analytically), point-triangle barrier blocks, affine-body 12x12
// blocks.  Output uses the reference's matrix contract
// (linear_system/linear_system/global_matrix.h:18-203): upper-triangular 3x3
// block triplets, unique and sorted by (row, col), each block 9 doubles in
// Eigen column-major order, global block order [4 rows per affine body][FEM
// vertices].  CPU only; deterministic for a given seed regardless of the
// number of threads (counter-based random numbers, colored assembly).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct masx_gen_params
{
    int      nx, ny, nz;            // FEM hex cells per axis (0 = no FEM part)
    double   spacing;               // cell size [m]
    double   young;                 // median Young's modulus [Pa]
    double   young_sigma;           // lognormal sigma of E per tet (0 = constant)
    double   poisson;
    double   density;               // [kg/m^3]
    double   dt;                    // [s]
    double   f_noise;               // F = I + f_noise * N(0,1) per tet
    int      n_contacts;            // random point-triangle pairs among surface vertices
    double   contact_stiffness;     // median barrier stiffness
    double   contact_sigma;         // lognormal sigma of the barrier stiffness
    int      n_bodies;              // affine bodies (12 DOF each)
    int      n_body_pairs;          // random body-body 12x12 couplings
    int      n_coupled_bodies;      // bodies coupled to one FEM surface vertex each
    double   body_shift;            // body diagonal = A^T A + body_shift * I
    uint64_t seed;
} masx_gen_params;

typedef struct masx_gen_sizes
{
    int n_vertices;   // FEM vertices
    int n_tets;
    int n_tris;       // boundary triangles
    int n_pairs;      // point-triangle pairs (int4 each)
    int n_bodies;
    int block_rows;   // 4 * n_bodies + n_vertices
    int fem_offset;   // 4 * n_bodies
    long long nb;     // upper 3x3 blocks
} masx_gen_sizes;

typedef struct masx_gen_system masx_gen_system;

enum masx_gen_field
{
    MASX_GEN_POSITIONS = 0,  // double[n_vertices*3]
    MASX_GEN_TETS      = 1,  // int[n_tets*4]      (FEM-local vertex ids)
    MASX_GEN_TRIS      = 2,  // int[n_tris*3]      (FEM-local vertex ids)
    MASX_GEN_PAIRS     = 3,  // int[n_pairs*4]     (global ids: fem_offset is NOT added;
                             //                     FEM-local ids, collision node offset 0)
    MASX_GEN_VALUES    = 4,  // double[nb*9]       (col-major blocks)
    MASX_GEN_ROWS      = 5,  // int[nb]
    MASX_GEN_COLS      = 6,  // int[nb]
    MASX_GEN_RHS       = 7,  // double[3*block_rows]
    MASX_GEN_DM_INV    = 8,  // double[n_tets*9]   (row-major rest-shape inverse)
    MASX_GEN_VOLUME    = 9,  // double[n_tets]
    MASX_GEN_DEFGRAD   = 10, // double[n_tets*9]   (row-major F per tet)
    MASX_GEN_MU        = 11, // double[n_tets]     (Lame mu per tet)
    MASX_GEN_LAMBDA    = 12, // double[n_tets]     (Lame lambda per tet)
    MASX_GEN_MASS      = 13  // double[n_vertices] (lumped vertex mass)
};

void masx_gen_default_params(masx_gen_params* p);
int  masx_gen_create(const masx_gen_params* p, masx_gen_system** out);
int  masx_gen_sizes_of(const masx_gen_system* s, masx_gen_sizes* out);
int  masx_gen_copy(const masx_gen_system* s, int field, void* dst);
void masx_gen_destroy(masx_gen_system* s);
const char* masx_gen_last_error(void);

// Element Hessian of the generator's energy for one tet (12x12 row-major,
// already multiplied by the rest volume, PSD-projected), with Lame (mu,
// lambda).  Used by tests and the adjoint's finite-difference check.
int masx_gen_element_hessian(const double* dm_inv, double volume, const double* F, double mu,
                             double lambda, double* H12);

#ifdef __cplusplus
}
#endif
