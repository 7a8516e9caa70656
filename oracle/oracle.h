/*
 * oracle.h -- C API of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (exastamp_amd/, libxsmd.so) never links or calls it.
 *
 * It restates exaStamp's CPU operators for the short-range force path in plain
 * C++20 + OpenMP, in FP64, with each function citing the reference file:line
 * it follows (paths relative to the exaStamp source tree).  Atoms are kept in
 * cell-sorted order like exaNBody's grid and loops run over cells with
 * schedule(dynamic), as the reference CPU path does.
 *
 * All arrays passed in/out through this API are indexed by particle id
 * (the order the caller supplied them in), never by the internal cell order.
 */
#ifndef XSMD_ORACLE_H
#define XSMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

enum { OR_POT_LJ = 0, OR_POT_EAM = 1, OR_POT_LJ_MULTI = 2, OR_POT_JOHNSON = 3 };
enum { OR_ENERGY = 1, OR_VIRIAL = 2 };

/* --- spline tables (eam_alloy.cpp:29-58) --------------------------------- */
/* f holds the n knot values f[1..n] (passed as f[0..n-1]); coeffs receives
 * (n+1) rows of 8 doubles, row 0 unused, exactly as EamAlloyTools::interpolate. */
int or_interpolate(int n, double delta, const double* f, double* coeffs);

/* --- setfl file (eam_alloy.cpp:66-201) ----------------------------------- */
int or_setfl_header(const char* path, int* nelements, int* nrho, double* drho,
                    int* nr, double* dr, double* rcut);
/* frho: nel*nrho, rhor: nel*nr, z2r: nel(nel+1)/2*nr, mass: nel */
int or_setfl_data(const char* path, double* frho, double* rhor, double* z2r, double* mass);

/* --- context ------------------------------------------------------------- */
or_ctx* or_create(void);
void or_destroy(or_ctx* c);
int or_set_threads(or_ctx* c, int nthreads);

int or_set_system(or_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                  const double* vx, const double* vy, const double* vz, const uint8_t* type,
                  const double lo[3], const double hi[3], const int periodic[3],
                  int nspecies, const double* masses);

int or_set_lj(or_ctx* c, double epsilon, double sigma, double rcut);
/* lj_multi_force parameters: one (epsilon, sigma, rcut) per unordered pair of the
 * ntypes species, entry b*(b+1)/2 + a for types a <= b. */
int or_set_lj_multi(or_ctx* c, int ntypes, const double* epsilon, const double* sigma, const double* rcut);

/* johnson_force: the 19 EamJohnsonParameters (johnson.h:30-51) in internal units,
 * order re fe rhoe alpha beta A B kappa lambda Fn0 Fn1 Fn2 Fn3 F0 F1 F2 F3 Fo eta. */
int or_set_johnson(or_ctx* c, const double* params, double rcut);
/* point values: out = rho(r), rho'(r), phi(r), phi'(r); F(rho), F'(rho) */
void or_johnson_eval(const double* params, double r, double out[4]);
void or_johnson_embed(const double* params, double rho, double* F, double* Fp);

/* EAM tables from knot values (setfl layout), splined by or_interpolate. */
int or_set_eam(or_ctx* c, int nel, int nrho, double drho, int nr, double dr, double rcut,
               const double* frho, const double* rhor, const double* z2r);

/* types slot of eam_alloy_force (eam_potential_multimat.cu:90,154-166): bit t set =
 * species t takes part; a pair counts when both of its species do, and an atom of a
 * disabled species gets rho = 0, no embedding term and no force.  ~0u = every species
 * (the empty list, the default). */
int or_set_eam_types(or_ctx* c, unsigned mask);

/* --- operators ----------------------------------------------------------- */
/* wrap + sort into cells + build the full neighbour list with rlist = rcut + skin */
int or_build_neighbors(or_ctx* c, double rcut, double skin);
/* zero_force_energy + force + force_to_accel */
int or_compute_forces(or_ctx* c, int potential, int flags);
/* a single pass of the EAM operator with the reference's flags */
int or_eam_pass(or_ctx* c, int eam_rho, int eam_rho2emb, int eam_force, int flags);
int or_zero_force_energy(or_ctx* c);
int or_force_to_accel(or_ctx* c);
int or_push_f_v_r(or_ctx* c, double dt);
/* push_f_v_r with its xform_mode: inv_xform = 1 (INV_XFORM) maps the step by the
 * inverse xform, 0 (IDENTITY) adds it as is. */
int or_push_f_v_r_mode(or_ctx* c, double dt, int inv_xform);
/* Domain xform (row-major grid -> physical map; NULL = identity).  Invalidates the
 * list.  -1 when det(m) <= 0. */
int or_set_xform(or_ctx* c, const double* m);
int or_push_f_v(or_ctx* c, double dt);
int or_displ_over(or_ctx* c, double threshold);
int or_thermo(or_ctx* c, double out[27]);

/* NVE loop as config_numerical_schemes.msp:44-52.  thermo_out (may be NULL)
 * receives (nsteps+1) x 27 doubles (step 0 = initial state).  Returns the
 * number of neighbour-list rebuilds (>= 1) or a negative error. */
int or_md_run(or_ctx* c, int potential, int nsteps, double dt, double rcut, double skin,
              double* thermo_out);
/* The steps alone, continuing from the current list and accelerations; thermo_out
 * (may be NULL) receives nsteps x 27.  Returns the number of rebuilds. */
int or_md_steps(or_ctx* c, int potential, int nsteps, double dt, double rcut, double skin,
                double* thermo_out);

/* --- read back (id order) ------------------------------------------------ */
int64_t or_n(const or_ctx* c);
int or_get_state(const or_ctx* c, double* x, double* y, double* z, double* vx, double* vy,
                 double* vz, double* ax, double* ay, double* az, double* ep, double* vir9,
                 double* rho_dEmb);
/* neighbour list by id: counts[n], then total entries; ids/imgs filled when non-NULL.
 * imgs holds 3 ints per entry.  Entries of atom i are in the oracle's list order. */
int64_t or_neighbor_count(const or_ctx* c, int64_t* counts);
int or_neighbor_get(const or_ctx* c, int64_t* ids, int32_t* imgs);
/* the internal cell order: perm[k] = id of the atom at sorted position k */
int or_get_order(const or_ctx* c, int64_t* perm);

/* brute-force O(N^2) neighbour sets with the same d2 contract, by id,
 * sorted by (id, image).  Small systems only. */
int64_t or_brute_neighbor_count(const or_ctx* c, double rlist, int64_t* counts);
int or_brute_neighbor_get(const or_ctx* c, double rlist, int64_t* ids, int32_t* imgs);

/* --- scalar spline evaluators (eam_alloy.h:198-313) ----------------------- */
double or_eam_rho(const or_ctx* c, double r, int itype, int jtype);
void or_eam_pair(const or_ctx* c, double r, double fpi, double fpj, int itype, int jtype,
                 double* fpair_over_r, double* phi);
void or_eam_embed(const or_ctx* c, double rho, int itype, double* F, double* Fp);
void or_lj_energy(const or_ctx* c, double r, double* e, double* de);

/* unit constants used (CODATA-2018) */
double or_eV(void);       /* 1 eV in Da.A^2/ps^2 */
double or_kB(void);       /* Boltzmann constant in Da.A^2/ps^2/K */

#ifdef __cplusplus
}
#endif

#endif
