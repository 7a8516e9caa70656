/*
 * xsmd.h -- C ABI of libxsmd.so, the MI355X (gfx950) short-range force path
 * that stands in for exaStamp's operators of the same names.
 *
 * Conventions
 *  - Every entry point returns an xs_status (0 = XS_OK) and never aborts; the
 *    message of the last failure on a context is xs_last_error(ctx).  The
 *    reference operators call fatal_error()/std::abort() instead
 *    (e.g. pair_potential_impl.hxx:135-147, eam_potential_multimat.cu:235-238).
 *  - Host pointers are plain host memory and are copied in/out.  The only
 *    device pointers are those named *_dev.
 *  - Per-atom host arrays are indexed by particle id = the position the atom
 *    had in the arrays given to xs_system_set.  Internally atoms live in
 *    cell-sorted order, like exaNBody's grid, and are re-sorted at every
 *    neighbour-list rebuild.
 *  - Units are exaStamp's internal units (include/exaStamp/unit_system.h:26-36):
 *    Angstrom, Dalton, picosecond; energy in Da.A^2/ps^2.
 *  - Calls on one context are ordered on the context's HIP stream and are not
 *    thread-safe.  Separate contexts may be used from separate threads.
 *  - One context holds one system made of nrep independent periodic boxes
 *    ("replicas"), used for batches of small configurations; nrep = 1 is the
 *    usual single box.
 */
#ifndef XSMD_H
#define XSMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xs_ctx xs_ctx;

typedef enum {
  XS_OK = 0,
  XS_E_ARG = 1,      /* invalid argument */
  XS_E_STATE = 2,    /* call out of order (no system, no list, no tables...) */
  XS_E_HIP = 3,      /* HIP runtime error */
  XS_E_ALLOC = 4,    /* device allocation failed */
  XS_E_IO = 5,       /* file could not be read or parsed */
  XS_E_LIMIT = 6     /* a size limit of this implementation was exceeded */
} xs_status;

/* compute flags */
#define XS_ENERGY 1u   /* per-atom potential energy ep */
#define XS_VIRIAL 2u   /* per-atom virial tensor (9 components, row-major) */
#define XS_MIXED 4u    /* FP32 pair/spline math; FP64 positions, displacements,
                          accumulation and reductions.  Without it: all FP64. */
#define XS_CONTINUE 8u /* xs_md_run: continue a previous run (valid list and current
                          accelerations): skip the initial build and force step */
#define XS_SYMMETRIC 16u /* Newton-symmetric evaluation (eam_symmetry: true,
                            lj_compute_force_symetric): each pair once, from a half of
                            the neighbour list, the partner's share added with FP64
                            atomics.  Operators only (not xs_md_run); the order of the
                            atomic sums, hence the last bits, varies from run to run. */

/* potentials for xs_md_run */
#define XS_POT_LJ 0
#define XS_POT_EAM 1

/* per-atom fields for xs_field_get */
#define XS_FIELD_EP 0        /* n doubles */
#define XS_FIELD_VIRIAL 1    /* 9n doubles */
#define XS_FIELD_RHO_DEMB 2  /* n doubles: rho after the density pass, F'(rho) after rho2emb */
#define XS_FIELD_ACCEL 3     /* 3n doubles (x then y then z blocks) */

int xs_version(void);

/* --- context --------------------------------------------------------------- */
/* Replaces the device set-up of hw_device_init/init_cuda (data/config/main-config.msp:87-91). */
xs_status xs_ctx_create(int device, xs_ctx** out);
xs_status xs_ctx_destroy(xs_ctx* ctx);
const char* xs_last_error(const xs_ctx* ctx);
/* Wait for all work queued on the context's stream. */
xs_status xs_synchronize(xs_ctx* ctx);
/* No counterpart in exaStamp.  Orders every later call of the context on `stream` (a
 * hipStream_t of the context's device, e.g. the caller's framework stream; NULL is the
 * device's null stream), so that device buffers the caller's own kernels write or read
 * (the *_dev arguments) need no further synchronisation.  Waits for the work already
 * queued on the current stream first.  The caller keeps `stream` alive while the
 * context uses it.  On such a stream the force operators do not wait for their block
 * passes' capacity flag (a brick beyond the capacity measured at the list build,
 * XS_E_STATE): the flag travels to the host with the work, and the first later call
 * that finds it arrived (xs_synchronize waits for it) returns XS_E_STATE. */
xs_status xs_ctx_set_stream(xs_ctx* ctx, void* stream);
/* Back to the context's own stream (the one xs_ctx_create made), after waiting for
 * the work queued on the current one. */
xs_status xs_ctx_own_stream(xs_ctx* ctx);
/* Free and total device memory of the context's GPU (hipMemGetInfo). */
xs_status xs_device_memory(xs_ctx* ctx, size_t* free_bytes, size_t* total_bytes);

/* --- species and system ------------------------------------------------------ */
/* Masses per atom type (ParticleSpecies, src/particle_species/include/exaStamp/
 * particle_species/particle_specie.h:52-122).  1 <= nspecies <= 7. */
xs_status xs_species_set(xs_ctx* ctx, int nspecies, const double* masses);

/* Load the atoms (replaces the grid filled by setup_system/lattice,
 * src/setup_system/bulk_lattice.cpp:17-31, and the rx..vz/type fields of
 * include/exaStamp/fields.h:64,69-79).  rep (may be NULL when nrep == 1) gives
 * the replica of each atom; lo/hi hold 3*nrep box bounds; periodic[3] applies
 * to all replicas.  Accelerations, energies and virials are zeroed. */
xs_status xs_system_set(xs_ctx* ctx, int64_t n, const double* rx, const double* ry,
                        const double* rz, const double* vx, const double* vy,
                        const double* vz, const uint8_t* type, const uint16_t* rep,
                        int nrep, const double* lo, const double* hi, const int periodic[3]);
int64_t xs_system_size(const xs_ctx* ctx);
/* Copy state back in id order; any pointer may be NULL. */
xs_status xs_system_get(xs_ctx* ctx, double* rx, double* ry, double* rz, double* vx,
                        double* vy, double* vz);
xs_status xs_field_get(xs_ctx* ctx, int field, double* out);
/* Domain xform of a deformed domain (exaNBody's Domain::xform / set_xform; read by
 * the pair operators at pair_potential_impl.hxx:388-413, which switch to the
 * LinearXForm traversal when it is not the identity, by the thermodynamic state for
 * the volume at simulation_thermodynamic_state.cpp:203-211, and by push_f_v_r's
 * INV_XFORM mode, push_vec3_2nd_order_xform_minimization.cpp:34).  m is row-major
 * 3x3 and maps grid coordinates (the positions given to xs_system_set and returned by
 * xs_system_get) to physical ones; NULL or the identity restores the plain form.
 * Pair displacements, forces, velocities and virials are physical.  det(m) must be
 * positive (XS_E_ARG otherwise).  The list becomes invalid: call xs_chunk_neighbors
 * before the next force operator.  Applies to every replica. */
xs_status xs_domain_xform(xs_ctx* ctx, const double m[9]);
/* Overwrite velocities (id order), e.g. for time-reversal tests. */
xs_status xs_velocity_set(xs_ctx* ctx, const double* vx, const double* vy, const double* vz);

/* --- neighbour list ------------------------------------------------------------ */
/* chunk_neighbors (exaNBody; slots data/config/config_move_particles.msp:54-68,
 * consumed at eam_potential_multimat.cu:94,180-181 and pair_potential_impl.hxx:111):
 * wrap atoms into the box, re-sort them by cell (move_particles), then build the
 * full per-atom list of every (atom, periodic image) with d^2 < (rcut_max+rcut_inc)^2.
 * The list is atom-granular, the chunk_size = 1 form; chunk_size must be a power of
 * two and any value gives identical interaction sets (a larger chunk only widens
 * the reference's list, never the set within rcut). */
xs_status xs_chunk_neighbors(xs_ctx* ctx, double rcut_max, double rcut_inc, int chunk_size);
/* particle_displ_over (config_move_particles.msp:19-23): *result = 1 when some atom
 * moved more than threshold since the last list build. */
xs_status xs_particle_displ_over(xs_ctx* ctx, double threshold, int* result);
/* Neighbour list in id space for checking: counts[n] (entries per atom); returns the
 * total through *total.  ids/imgs (3 ints per entry) may be NULL; entries of one atom
 * are in list order, atoms in id order. */
xs_status xs_neighbors_get(xs_ctx* ctx, int64_t* counts, int64_t* total, int64_t* ids,
                           int32_t* imgs);
/* perm[k] = id of the atom at internal (cell-sorted) position k. */
xs_status xs_order_get(xs_ctx* ctx, int64_t* perm);
/* Neighbour statistics (the chunk_neighbors_stats diagnostic of
 * data/samples/perf/microjet_lj.msp:25-28): list entries and entries with
 * d^2 <= rcut^2 at the current positions. */
xs_status xs_pair_count(xs_ctx* ctx, double rcut, int64_t* entries, int64_t* within_rcut);
/* Which list the last build produced (diagnostic; no counterpart in the reference):
 * *kind = 0 for the global (index, image) list, 1 for the cell-block list of the
 * LDS-staged EAM kernels; *blocks = bricks of cells, *halo_capacity = LDS slots per
 * brick, *central_max = most central atoms in one brick, brick[0..2] = cells per
 * brick along x, y, z (all 0 for the global list; any pointer may be NULL).  The
 * environment variable XSMD_LIST=global, read when a context is created, keeps every
 * build on the global list; XSMD_BLOCK_ATOMS=n bounds the expected central atoms
 * per brick (default 230). */
xs_status xs_list_info(xs_ctx* ctx, int* kind, int64_t* blocks, int* halo_capacity, int* central_max, int* brick);

/* --- operators -------------------------------------------------------------------- */
/* zero_force_energy (src/compute/zero_force_energy.cu:38-147): a = ep = virial = 0. */
xs_status xs_zero_force_energy(xs_ctx* ctx);
/* force_to_accel (src/compute/force_to_accel.cu:38-122): a /= mass[type]. */
xs_status xs_force_to_accel(xs_ctx* ctx);
/* lj_compute_force (pair_potential_impl.hxx:39-513 with lennard_jones.h:33-50):
 * adds LJ forces, shifted energies (e - E(rcut)) and virials to the atoms. */
xs_status xs_lj_compute_force(xs_ctx* ctx, double epsilon, double sigma, double rcut,
                              unsigned flags);
/* The parameters slot of lj_compute_force alone (LennardJonesParms, lennard_jones.h:33-37,
 * decoded at :60-71), for xs_md_run. */
xs_status xs_lj_parameters(xs_ctx* ctx, double epsilon, double sigma, double rcut);
/* lj_multi_force (pair_potential_impl_multiparam.cu with the per-type-pair table of
 * pair_potential_force_op_multiparam.h:58-180 and pair_potential_impl.hxx:282-320):
 * LJ with its own epsilon, sigma and cutoff for every unordered pair of the ntypes
 * species (ntypes = the species count), entry b*(b+1)/2 + a for types a <= b; each
 * pair's energy is shifted by its E(rcut), a pair counts when r <= its rcut, and
 * epsilon 0 gives no interaction.  Operator form (adds forces, energies, virials). */
xs_status xs_lj_multi_force(xs_ctx* ctx, int ntypes, const double* epsilon, const double* sigma,
                            const double* rcut, unsigned flags);
/* eam_alloy_init + parameters (eam_alloy.cpp:66-278): tabulated knots in setfl
 * layout.  frho: nel*nrho values, rhor: nel*nr, z2r: nel(nel+1)/2*nr (packed lower
 * triangle 11,21,22,31,...), z2r = r*phi in eV.A, F in eV. */
xs_status xs_eam_alloy_tables(xs_ctx* ctx, int nel, int nrho, double drho, int nr, double dr,
                              double rcut, const double* frho, const double* rhor,
                              const double* z2r);
/* Same from a setfl (.eam.alloy) file; masses (nullable, >= 7 doubles) and the
 * element count are returned. */
xs_status xs_eam_alloy_load(xs_ctx* ctx, const char* path, double* masses, int* nel);
/* eam_alloy_force (eam_potential_multimat.cu:65-259) with its eam_rho / eam_rho2emb /
 * eam_force slots; eam_ghost is not needed (no ghosts); eam_symmetry is XS_SYMMETRIC
 * (SymRhoOp / SymForceOp with NewtonSym, eam_force_op_multimat.h:114-343). */
xs_status xs_eam_alloy_force(xs_ctx* ctx, double rcut, int eam_rho, int eam_rho2emb,
                             int eam_force, unsigned flags);
/* The types slot of eam_alloy_force (eam_potential_multimat.cu:90, applied at
 * :154-166 through m_pair_enabled and checked at eam_force_op_multimat.h:149-153,
 * 180, 191, 303-308): bit t of `enabled` set = species t takes part.  A pair
 * contributes when both of its species take part; an atom of a species that does not
 * keeps rho = 0 and F'(rho) = 0, gets no embedding energy and no force.  ~0u (the
 * default) = every species, the empty list.  Applies to every later
 * xs_eam_alloy_force, xs_md_run and xs_eam_adjoint on the context. */
xs_status xs_eam_alloy_types(xs_ctx* ctx, unsigned enabled);
/* FP32 tables in s = r^2 that the mixed single-element force pass reads instead of
 * the knot rows (no square root or division per pair): *s_pieces = cubic pieces per
 * table (0 = none: the force pass reads the FP32 knot rows; also with the environment
 * variable XSMD_EAM_S_TABLE=0 at context creation), *r_lo = the shortest distance
 * they cover (closer pairs read the knot rows), err[3] = their largest FP32 errors
 * against the FP64 spline in FP32 ulps of the largest |rho'/r|, |phi| and |phi'/r|
 * over the range (taken when the slopes stay within 8 and phi within 4).  No
 * counterpart in exaStamp (a layout of the mixed form); any output may be NULL. */
xs_status xs_eam_table_info(xs_ctx* ctx, int* s_pieces, double* r_lo, double err[3]);
/* johnson_force (single-material EAM template eam_potential.cu:62-193 with the
 * analytic functions of johnson.h:59-167): both passes, FP64, full list.
 * params: the 19 EamJohnsonParameters in internal units, in the order
 *   re fe rhoe alpha beta A B kappa lambda Fn0 Fn1 Fn2 Fn3 F0 F1 F2 F3 Fo eta.
 * johnson_emb runs EmbOp (adds F(rho_i) to ep, keeps F'(rho_i)); johnson_force runs
 * ForceOp (adds the force, sum phi/2 to ep, the virial with XS_VIRIAL): both = the
 * johnson_force operator, emb only = johnson_emb, force only = johnson_force_reuse_emb
 * (XS_E_STATE unless johnson_emb ran at the current atom order).  Energy is always
 * accumulated, as EmbOp and ForceOp do.  XS_MIXED and XS_SYMMETRIC are rejected.
 * Leaves no eam_alloy rho/F' state. */
xs_status xs_johnson_force(xs_ctx* ctx, const double params[19], double rcut, int johnson_emb,
                           int johnson_force, unsigned flags);
/* push_f_v_r (config_numerical_schemes.msp:23-25): r += v*dt + a*dt^2/2,
 * dt = dt * dt_scale; xform_mode XS_XFORM_INV (the scheme's INV_XFORM) maps the
 * physical step into grid space by the inverse domain xform first
 * (push_vec3_2nd_order_xform_minimization.cpp:34), XS_XFORM_IDENTITY adds it as is;
 * the two agree when the xform is the identity. */
#define XS_XFORM_IDENTITY 0
#define XS_XFORM_INV 1
xs_status xs_push_f_v_r(xs_ctx* ctx, double dt, double dt_scale, int xform_mode);
/* push_f_v (config_numerical_schemes.msp:24-28): v += a*dt, dt = dt * dt_scale. */
xs_status xs_push_f_v(xs_ctx* ctx, double dt, double dt_scale);
/* simulation_thermodynamic_state reduction (src/thermo_state/
 * simulation_thermodynamic_state.cpp:101-163): out[27] = virial(9), kinetic tensor
 * (9), momentum(3), kinetic energy(3), potential energy, mass, particle count. */
xs_status xs_thermo(xs_ctx* ctx, double out27[27]);
/* No counterpart in exaStamp.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops): xs_thermo with the same kernels and
 * the same summation order, but the partial sums go to a fixed device array of
 * 1024 x 26 doubles (208 KiB, one per device, held until the process ends) and the
 * 27 results are written to device memory at out27_dev.  The call allocates nothing,
 * copies nothing to the host and does not synchronise: read out27_dev after
 * xs_synchronize (or after ordering another stream behind the context's).  Every
 * partial sum it reads is written by the same call, so nothing needs clearing; calls
 * that overlap on one device (several contexts on other streams or threads) race on
 * the array and must be serialised by the caller. */
xs_status xs_thermo_noalloc(xs_ctx* ctx, double* out27_dev);

/* --- EAM adjoint --------------------------------------------------------------------- */
/* No counterpart in exaStamp: reverse mode of eam_alloy_force with respect to the
 * tabulated knots, for fitting tables to reference forces and energies.  Call after
 * xs_eam_alloy_force with eam_rho, eam_rho2emb and eam_force in the same call (its
 * rho_i and F'(rho_i) are the checkpoint; they are not recomputed).
 *   dL_dF : 3n doubles, dL/dF for each atom's force (id order, x y z per atom),
 *           forces in internal units (Da.A/ps^2)
 *   dL_dE : nrep doubles, dL/dE for each replica's total potential energy
 *   grad  : receives dL/dtheta for the knots in setfl order: F tables (nel*nrho),
 *           rho tables (nel*nr), r*phi tables (nel(nel+1)/2*nr), each per file unit
 *           (eV, density, eV.A).
 * flags: XS_MIXED reads the FP32 copies of the tables (distances, knot fractions,
 * weights and accumulation stay FP64).  The checkpoint may come from either
 * precision of the forward. */
xs_status xs_eam_adjoint(xs_ctx* ctx, double rcut, const double* dL_dF, const double* dL_dE,
                         double* grad, unsigned flags);
/* xs_eam_adjoint with its three arrays in device memory, for the incoming gradients
 * dL_dF_dev and dL_dE_dev times `scale` (e.g. 1/EV for gradients with respect to forces
 * and energies in eV units): nothing is copied to or from the host and the call does not
 * synchronise (grad_dev is complete when the context's stream reaches it).  No
 * counterpart in exaStamp; for fitting loops that keep the knots, the forces and the
 * loss on the device. */
xs_status xs_eam_adjoint_dev(xs_ctx* ctx, double rcut, const double* dL_dF_dev, const double* dL_dE_dev,
                             double scale, double* grad_dev, unsigned flags);
/* No counterpart in exaStamp (the fitting counterpart of eam_alloy_init for new knot
 * values): theta_dev holds knots in xs_eam_adjoint's setfl order, on the grid (nel,
 * nrho, drho, nr, dr, rcut) of the tables loaded last, in device memory.  The spline
 * coefficients of both precisions are built on the device (the same construction and
 * bits as xs_eam_alloy_tables); the FP32 s = r^2 tables are dropped, so the mixed
 * force pass reads the FP32 knot rows until the next xs_eam_alloy_tables.  No copy to
 * or from the host, no synchronisation. */
xs_status xs_eam_theta_dev(xs_ctx* ctx, const double* theta_dev);
/* No counterpart in exaStamp (the device form of reading forces and energies back for
 * a fitting loss): F_dev (3n doubles, id order, x y z per atom) = m a_i of the current
 * accelerations (valid after force_to_accel), E_dev (nrep doubles) = each replica's
 * total potential energy (sum of ep, needs a force call with XS_ENERGY), both in
 * internal units times `scale` (e.g. 1/EV for eV units).  No copy to the host, no
 * synchronisation. */
xs_status xs_forces_energies_dev(xs_ctx* ctx, double* F_dev, double* E_dev, double scale);
/* No counterpart in exaStamp: reverse mode of lj_compute_force (pair_potential_impl.hxx:
 * 39-513 with lennard_jones.h:40-50) with respect to its parameters, for fitting
 * epsilon and sigma to reference forces and energies.  Differentiates the forces
 * F_i = sum_j e'(r)/r (x_j - x_i) and the shifted pair energies 1/2 (e(r) - e(rcut))
 * per list entry, the shift's own dependence on epsilon and sigma included
 * (energy_cutoff, pair_potential_impl.hxx:487-498).  Needs the neighbour list of the
 * current positions (those of the forward call); one FP64 sweep of the list.
 *   dL_dF : 3n doubles, dL/dF per atom (id order, x y z), forces in internal units
 *   dL_dE : nrep doubles, dL/dE per replica's total potential energy
 *   grad  : receives dL/depsilon (epsilon in internal energy units) and dL/dsigma
 *           (sigma in A)
 * flags must be 0. */
xs_status xs_lj_adjoint(xs_ctx* ctx, double epsilon, double sigma, double rcut, const double* dL_dF,
                        const double* dL_dE, double grad[2], unsigned flags);
/* Knot-gradient of the spline construction alone (transpose of the knot -> c3..c6
 * map): G has (n+1) rows of 4 coefficient gradients, out n knots.  Host only. */
xs_status xs_spline_adjoint(int n, double delta, const double* G, double* out);

/* --- fused NVE loop ---------------------------------------------------------------- */
typedef struct xs_run_stats {
  int64_t steps;
  int64_t rebuilds;     /* neighbour-list builds including the first */
  int64_t max_neighbors;
  double seconds;       /* wall time of the loop (GPU synchronised) */
} xs_run_stats;

/* verlet_nve (config_numerical_schemes.msp:44-52) for nsteps steps, starting with
 * first_iteration's force computation (main-config.msp:145-152), rebuilding the list
 * when an atom moved more than skin/2.  The potential was set by a prior
 * xs_lj_compute_force / xs_eam_alloy_tables call.  flags: XS_MIXED selects the
 * precision; when thermo_every > 0, energies and virials are computed on every
 * thermo_every-th step (and step 0) and the 27-double state is written to
 * thermo_out[(step/thermo_every)*27].  stats may be NULL. */
xs_status xs_md_run(xs_ctx* ctx, int potential, int nsteps, double dt, double rcut,
                    double skin, unsigned flags, int thermo_every, double* thermo_out,
                    xs_run_stats* stats);

/* --- timing ------------------------------------------------------------------------ */
/* GPU time per kernel class, measured with HIP events recorded on the context's stream
 * around every launch of that class (the per-operator timers of onika's profiling
 * mode, data/config/config_defaults.msp:17-23). */
#define XS_PROF_NEIGHBORS 0   /* whole list rebuild: wrap, sort, permute, build */
#define XS_PROF_EAM_RHO 1     /* density pass (+ embedding) */
#define XS_PROF_EAM_FORCE 2   /* force pass (+ force_to_accel and half kick when fused) */
#define XS_PROF_LJ 3
#define XS_PROF_DRIFT 4       /* push_f_v_r + push_f_v + displacement check */
#define XS_PROF_THERMO 5
#define XS_PROF_ADJOINT 6     /* both EAM adjoint sweeps */
#define XS_PROF_NCLASSES 7
/* Streaming copy of `bytes` between two device buffers (allocated and freed by the
 * call), repeated `reps` times; *gbs receives (read + write bytes) / time of the
 * fastest repeat.  The achievable-bandwidth reference for the roofline, and a
 * known-size kernel (k_stream_copy) for calibrating HBM byte counters. */
xs_status xs_stream_copy(xs_ctx* ctx, size_t bytes, int reps, double* gbs);
/* Gather ceiling of the mixed single-element EAM passes on the current cell-block
 * list and tables: average ms over `reps` launches of a kernel that does everything
 * the density pass (force_pass = 0: knot rows) or the force pass (force_pass = 1:
 * s-table rows) does to reach its table rows -- staging, list walk, neighbour reads,
 * FP64 displacement, cutoff test, row index, the row gathers themselves -- and none of
 * the spline or pair arithmetic (level 0).  Its time over the pass's own is the share
 * of the pass that its data movement accounts for (the roofline of a gather-bound
 * pass).  Levels 1-3 take parts away to split that time: 1 without the table-row
 * gathers, 2 the list words only (no neighbour records or displacements), 3 the
 * staging of the bricks' halos only.  Writes no simulation state. */
xs_status xs_gather_ceiling(xs_ctx* ctx, double rcut, int force_pass, int level, int reps, double* ms);
/* on != 0 enables and clears the counters; on == 0 disables. */
xs_status xs_profile_enable(xs_ctx* ctx, int on);
/* ms_total[XS_PROF_NCLASSES] and launches[XS_PROF_NCLASSES] (either may be NULL). */
xs_status xs_profile_get(xs_ctx* ctx, double* ms_total, int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* XSMD_H */
