/*
 * pfem_oracle.h -- CPU oracle for the Newton-step hot path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (vinaysrao1/plasticity, lagrangian/src) used as the checker for the
 * HIP product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links or calls it.
 *
 * Conventions (all follow the reference):
 *   3x3 matrices     column-major, m[i + 3*j] = M_ij          (StaticArrays layout)
 *   Voigt 6-vectors  [xx, yy, zz, xy, yz, zx]; strain = engineering shear,
 *                    stress / Cp_inv = physical shear        (Materials.jl:18-19)
 *   element arrays   node-major: Xe[3*a + d], ue[3*a + i]     (Assembly.jl:51-57)
 *   Ke               24x24 column-major, Ke[r + 24*c]          (Elements.jl:427)
 *   A = dP/dF        9x9 column-major, F index (j-1)*3+i      (FiniteStrain.jl:494)
 *   Gauss state      6 x ngp column-major, st[6*gp + k], gp = 8*e + g
 *   global indices   0-based here (the reference is 1-based)
 */
#ifndef PFEM_ORACLE_H
#define PFEM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_mat {
    double E, nu, sy0, Hiso, Hkin, ssat, delta; /* inputs */
    double G, K, lam;                            /* derived */
    double C[36];                                /* row-major 6x6 Voigt C */
} or_mat;

/* J2Material constructor with the reference's validity guards
 * (Materials.jl:82-108).  Returns 0, or k>0 naming the guard that failed. */
int or_material_init(or_mat *m, double E, double nu, double sy0, double Hiso,
                     double Hkin, double ssat, double delta);

/* return_map (Materials.jl:159-250).  D is row-major 6x6. */
void or_return_map(const or_mat *m, const double eps[6], const double epn[6],
                   const double bn[6], double an, double sig[6], double ep[6],
                   double bt[6], double *aout, double D[36]);

/* Symmetric 3x3 eigen-decomposition (cyclic Jacobi), ascending eigenvalues,
 * eigenvectors as columns of V (column-major). */
void or_eig_sym3(const double A[9], double w[3], double V[9]);

typedef struct or_kin {
    double ee[6];   /* trial elastic Hencky strain (eng. shear) */
    double b[3];    /* eigenvalues of be_tr */
    double n[9];    /* eigenvectors (columns) */
    double Finv[9];
    double J;
    int ok;
} or_kin;

void or_finite_kinematics(const double F[9], const double cpi[6], or_kin *k);

/* polar_RU (FiniteStrain.jl:190-205) and dR_polar (:215-228); 3x3 column-major. */
void or_polar_RU(const double F[9], double R[9], double U[9]);
void or_dR_polar(const double R[9], const double U[9], const double dF[9], double dR[9]);
/* dtau_dbeta (FiniteStrain.jl:245-262), row-major 6x6. */
void or_dtau_dbeta(const or_mat *m, const double ee[6], const double bsp[6], double an,
                   double out[36]);

/* finite_stress_update (FiniteStrain.jl:292-346).  bref_n may be NULL (zero back
 * stress); bref_new, cpi_new, bsp_out, R_out, U_out may be NULL. */
void or_finite_stress_update(const or_mat *m, const or_kin *k, const double F[9], const double epn[6],
                             const double bref_n[6], double an, double tau[6], double ep[6],
                             double bref_new[6], double *aout, double D[36], double cpi_new[6],
                             double bsp_out[6], double R_out[9], double U_out[9]);

/* dPdF (FiniteStrain.jl:503-547).  A column-major 9x9.  bref, R, U, dtdb NULL: the
 * Hkin == 0 form without the back-stress rotation term. */
void or_dPdF(const or_kin *k, const double cpi[6], const double D[36],
             const double tau[6], const double F[9], const double bref[6], const double R[9],
             const double U[9], const double dtdb[36], double A[81]);

/* hex8 geometry (Elements.jl:62-75, 155-161, 517-519) */
void or_hex8_dshape(const double xi[3], double dN[24]);
void or_ref_grads(const double Xe[24], const double xi[3], double dNdX[24], double *detJ);

/* element_force_tangent_finite! (Elements.jl:405-494).
 * fbar: 0 = Hex8Finite, 1 = Hex8FiniteFbar.  State arrays hold the 8 GPs of
 * this element (6x8 / 8); beta is the back stress in the reference
 * configuration.  Returns 0, or 1 on an inverted GP (fail_g, fail_J filled,
 * nothing further written for that GP). */
int or_element_finite(int fbar, const or_mat *m, const double Xe[24],
                      const double ue[24], double *epsp, double *beta, double *abar,
                      double *cpinv, double *sig, int commit, double Fe[24],
                      double Ke[576], int *fail_g, double *fail_J);

/* element_force_tangent! (Elements.jl:323-360), small strain. */
void or_element_small(const or_mat *m, const double Xe[24], const double ue[24], double *epsp,
                      double *beta, double *abar, double *sig, int commit, double Fe[24],
                      double Ke[576]);

/* ---------------- mesh / pattern / colours ---------------- */

/* box_mesh (Mesh.jl:30-59) plus the seeded interior jitter (+-jit*h per axis,
 * splitmix64 counter RNG).  nodes 3 x nnodes, conn 8 x nelem (0-based). */
void or_box_mesh(double lx, double ly, double lz, int nx, int ny, int nz,
                 double jit, uint64_t seed, double *nodes, int64_t *conn);

uint64_t or_splitmix64(uint64_t x);
double or_jitter_u01(uint64_t seed, int64_t node, int comp);

/* Node adjacency + CSC skeleton (Assembly.jl:62-162).  Two-phase: first call
 * with rowval == NULL returns nnz via *nnz_out and fills colptr. */
int or_build_csc(int64_t nnodes, int64_t nelem, const int64_t *conn,
                 int64_t *colptr, int64_t *rowval, int64_t *nnz_out);

/* Greedy first-fit element colouring in element order; colour[e] in 0..63.
 * Returns the number of colours. */
int or_greedy_colors(int64_t nnodes, int64_t nelem, const int64_t *conn, int *color);

/* Colour-ordered assembly (Assembly.jl:370-416): K.nzval and R overwritten.
 * kind: 0 Hex8Finite, 1 Hex8FiniteFbar, 2 Hex8Small.
 * colour_order lists elements colour by colour; colour_ptr has ncolors+1
 * offsets.  nthreads <= 0 => OpenMP default.  Returns 0 or 1 (inversion).
 * mats: one material, or with matid (per element, may be NULL) an array of
 * them -- the return map is local to each Gauss point, so a two-phase block is
 * the same loop with the element's own material. */
int or_assemble(int kind, const or_mat *mats, const int32_t *matid, int64_t nnodes, int64_t nelem,
                const double *nodes, const int64_t *conn, const int64_t *colptr,
                const int64_t *rowval, int ncolors, const int64_t *colour_ptr,
                const int64_t *colour_order, const double *U, double *epsp, double *beta,
                double *abar, double *cpinv, double *sig, int commit,
                double *nzval, double *R, int nthreads, int64_t *fail_e,
                int *fail_g, double *fail_J);

/* impose_dirichlet! (BoundaryConditions.jl:62-109) on CSC (K, R). */
void or_impose_dirichlet(int64_t n, const int64_t *colptr, const int64_t *rowval,
                         double *nzval, double *R, int64_t nbc, const int64_t *dofs,
                         const double *values, int ramp, double lambda,
                         const double *U);

/* gauss_stress (Model.jl:188-208): sig = tau / det F (un-barred F at each Gauss point);
 * kind 2 (small strain) copies tau. */
void or_gauss_stress(int kind, int64_t nelem, const double *nodes, const int64_t *conn,
                     const double *U, const double *tau, double *sig);

#ifdef __cplusplus
}
#endif
#endif
