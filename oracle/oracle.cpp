// oracle.cpp -- FP64 CPU restatement of exaStamp's short-range force path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the GPU path (see oracle.h).
// Every function names the reference lines it restates; paths are relative to
// the exaStamp source tree.  exaNBody (grid, chunk_neighbors,
// compute_cell_particle_pairs, push_f_v_r/push_f_v) is not part of that tree;
// where it is restated, the exaStamp call site or deck that fixes its behaviour
// is cited instead.
//
// Build: see oracle/Makefile (g++ -O2 -fopenmp -mfma -ffp-contract=off).

#include "oracle.h"
#include "../include/xsmd_pair_geometry.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <string>
#include <vector>
#include <omp.h>

namespace {

// Exact SI charge and Boltzmann constants; CODATA-2022 atomic mass constant (the
// value the reference's recorded AlCu run implies through its density column,
// data/regression_new/potentials/eam/eam_alloy/thermodynamic_state.csv).
// exaStamp's internal units are Angstrom, Dalton,
// picosecond (include/exaStamp/unit_system.h:26-36), so the internal energy
// unit is 1 Da.A^2/ps^2 = amu * 1e4 J (print_thermodynamic_state.cpp:222-224).
constexpr double kAmu = 1.66053906892e-27;
constexpr double kElementaryCharge = 1.602176634e-19;
constexpr double kBoltzmann = 1.380649e-23;
constexpr double kEnergyUnitJ = kAmu * 1.0e4;
constexpr double kEV = kElementaryCharge / kEnergyUnitJ;
constexpr double kKB = kBoltzmann / kEnergyUnitJ;

// --------------------------------------------------------------------------
// Spline build: EamAlloyTools::interpolate, eam_alloy.cpp:29-58.  Row m
// (1-based) holds the cubic value coefficients in [3..6] and the quadratic
// derivative coefficients (already divided by delta) in [0..2]; rows are
// padded to 8 doubles like SplineCoeffs (eam_alloy.h:37-42).
// --------------------------------------------------------------------------
void interpolate(int n, double delta, const double* f, double* s)
{
  auto S = [&](int m, int k) -> double& { return s[8 * (size_t)m + k]; };
  std::fill(s, s + 8 * (size_t)(n + 1), 0.0);
  for (int m = 1; m <= n; m++) S(m, 6) = f[m - 1];

  S(1, 5) = S(2, 6) - S(1, 6);
  S(2, 5) = 0.5 * (S(3, 6) - S(1, 6));
  S(n - 1, 5) = 0.5 * (S(n, 6) - S(n - 2, 6));
  S(n, 5) = S(n, 6) - S(n - 1, 6);

  for (int m = 3; m <= n - 2; m++)
    S(m, 5) = ((S(m - 2, 6) - S(m + 2, 6)) + 8.0 * (S(m + 1, 6) - S(m - 1, 6))) / 12.0;

  for (int m = 1; m <= n - 1; m++) {
    S(m, 4) = 3.0 * (S(m + 1, 6) - S(m, 6)) - 2.0 * S(m, 5) - S(m + 1, 5);
    S(m, 3) = S(m, 5) + S(m + 1, 5) - 2.0 * (S(m + 1, 6) - S(m, 6));
  }
  S(n, 4) = 0.0;
  S(n, 3) = 0.0;

  for (int m = 1; m <= n; m++) {
    S(m, 2) = S(m, 5) / delta;
    S(m, 1) = 2.0 * S(m, 4) / delta;
    S(m, 0) = 3.0 * S(m, 3) / delta;
  }
}

struct Spline {
  int n = 0;
  std::vector<double> c;
  const double* row(int m) const { return &c[8 * (size_t)m]; }
};

// EamAlloyParameters + EamAlloyParametersRO (eam_alloy.h:44-196).
struct Eam {
  int nel = 0, nr = 0, nrho = 0, n_types = 0;
  double rdr = 0, rdrho = 0, rc = 0, rhomax = 0;
  std::vector<Spline> frho, rhor, z2r;
  int type2rhor[8][8] = {};
  int type2z2r[8][8] = {};
  int type2frho[8] = {};

  // initialize_types_table, eam_alloy.h:70-133 (1-based types, identity map).
  void init_types(int nt)
  {
    n_types = nt;
    int map[8];
    for (int i = 0; i < 8; i++) map[i] = -1;
    for (int i = 1; i <= nt; i++) map[i] = i - 1;
    const int nf = (int)frho.size();
    for (int i = 1; i <= nt; i++) type2frho[i] = (map[i] >= 0) ? map[i] : nf - 1;
    for (int i = 1; i <= nt; i++)
      for (int j = 1; j <= nt; j++) type2rhor[i][j] = map[i];
    for (int i = 1; i <= nt; i++)
      for (int j = 1; j <= nt; j++) {
        int irow = map[i], icol = map[j];
        if (irow == -1 || icol == -1) { type2z2r[i][j] = 0; continue; }
        if (irow < icol) { irow = map[j]; icol = map[i]; }
        int n = 0;
        for (int m = 0; m < irow; m++) n += m + 1;
        n += icol;
        type2z2r[i][j] = n;
      }
  }
};

// eam_alloy_rho_noderiv, eam_alloy.h:198-220
double eam_rho_noderiv(const Eam& e, double r, int itype, int jtype)
{
  itype += 1;
  jtype += 1;
  double p = r * e.rdr + 1.0;
  int m = (int)p;
  m = std::min(m, e.nr - 1);
  p -= m;
  p = std::min(p, 1.0);
  const double* c = e.rhor[e.type2rhor[itype][jtype]].row(m);
  return ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
}

// eam_alloy_mm_force, eam_alloy.h:222-265.  dr_fe is rj-ri on entry and
// (rj-ri)*fpair on exit.
void eam_mm_force(const Eam& e, double dr_fe[3], double& phi, double r, double fpi, double fpj,
                  int itype, int jtype)
{
  itype += 1;
  jtype += 1;
  double p = r * e.rdr + 1.0;
  int m = (int)p;
  m = std::min(m, e.nr - 1);
  p -= m;
  p = std::min(p, 1.0);

  const double* cij = e.rhor[e.type2rhor[itype][jtype]].row(m);
  const double rhoip = (cij[0] * p + cij[1]) * p + cij[2];
  const double* cji = e.rhor[e.type2rhor[jtype][itype]].row(m);
  const double rhojp = (cji[0] * p + cji[1]) * p + cji[2];
  const double* c = e.z2r[e.type2z2r[itype][jtype]].row(m);
  const double z2p = (c[0] * p + c[1]) * p + c[2];
  const double z2 = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
  const double recip = 1.0 / r;

  phi = z2 * recip;
  const double phip = (z2p * recip - phi * recip) * kEV;  // conversion_z2r = 1 eV.A
  phi *= kEV;
  const double psip = fpi * rhojp + fpj * rhoip + phip;
  const double fpair = psip * recip;
  dr_fe[0] *= fpair;
  dr_fe[1] *= fpair;
  dr_fe[2] *= fpair;
}

// eam_alloy_fEmbed_deriv, eam_alloy.h:267-287
double eam_fembed_deriv(const Eam& e, double rho, int itype)
{
  itype += 1;
  double p = rho * e.rdrho + 1.0;
  int m = (int)p;
  m = std::max(1, std::min(m, e.nrho - 1));
  p -= m;
  p = std::min(p, 1.0);
  const double* c = e.frho[e.type2frho[itype]].row(m);
  return ((c[0] * p + c[1]) * p + c[2]) * kEV;  // conversion_frho = 1 eV
}

// eam_alloy_fEmbed, eam_alloy.h:289-313 (linear extrapolation above rhomax)
void eam_fembed(const Eam& e, double rho, double& phi, double& fp, int itype)
{
  itype += 1;
  double p = rho * e.rdrho + 1.0;
  int m = (int)p;
  m = std::max(1, std::min(m, e.nrho - 1));
  p -= m;
  p = std::min(p, 1.0);
  const double* c = e.frho[e.type2frho[itype]].row(m);
  fp = (c[0] * p + c[1]) * p + c[2];
  phi = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
  if (rho > e.rhomax) phi += fp * (rho - e.rhomax);
  phi *= kEV;
  fp *= kEV;
}

// lj_compute_energy, lennard_jones.h:40-50
void lj_compute_energy(double eps, double sigma, double r, double& e, double& de)
{
  double ratio = sigma / r;
  double ratio2 = ratio * ratio;
  double ratio6 = ratio2 * ratio2 * ratio2;
  double ratio12 = ratio6 * ratio6;
  e = 4. * eps * (ratio12 - ratio6);
  de = (-24. * eps * (2. * ratio12 - ratio6)) / r;
}

// onika tensor(u,v): m_ab = u_a * v_b, row-major m11 m12 m13 m21 ... m33
inline void add_tensor_scaled(double* vir, const double u[3], const double v[3], double s)
{
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) vir[3 * a + b] += (u[a] * v[b]) * s;
}

}  // namespace

// Johnson parameter indices, in the EamJohnsonParameters order (johnson.h:30-51)
enum { J_RE, J_FE, J_RHOE, J_ALPHA, J_BETA, J_A, J_B, J_KAPPA, J_LAMBDA, J_FN0, J_FN1, J_FN2, J_FN3, J_F0, J_F1,
       J_F2, J_F3, J_FO, J_ETA };

// ==========================================================================
// Context: exaNBody grid restated as cell-sorted flat SoA arrays.
// ==========================================================================
struct or_ctx {
  int nthreads = 0;
  int64_t n = 0;
  std::vector<double> x, y, z, vx, vy, vz, ax, ay, az, ep, vir, rho, xr, yr, zr;
  std::vector<uint8_t> type;
  std::vector<int64_t> id;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, L[3] = {0, 0, 0};
  int periodic[3] = {1, 1, 1};
  std::vector<double> mass;
  // Domain xform (grid -> physical, row-major) and its inverse; positions are grid
  // coordinates, pair displacements, forces and velocities physical.
  bool xf_on = false;
  double xf[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, xfi[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

  bool has_lj = false;
  double lj_eps = 0, lj_sigma = 0, lj_rcut = 0, lj_ecut = 0;
  // lj_multi_force: per type pair (packed triangle) epsilon, sigma, rcut, ecut
  int ljm_ntypes = 0;
  std::vector<double> ljm_eps, ljm_sigma, ljm_rcut, ljm_ecut;
  bool has_eam = false;
  Eam eam;
  // types slot of eam_alloy_force as per-species enable bits: a pair takes part when
  // both species are enabled (m_pair_enabled, eam_potential_multimat.cu:154-166)
  uint32_t eam_types = ~0u;
  bool eam_on(int t) const { return (eam_types >> t) & 1u; }
  // johnson_force: EamJohnsonParameters (johnson.h:30-51), internal units
  bool has_johnson = false;
  double jp[19] = {};
  double johnson_rc = 0;

  int nc[3] = {1, 1, 1};
  double inv_cs[3] = {0, 0, 0};
  std::vector<int64_t> cell_start;  // ncells+1
  std::vector<int64_t> nb_off;      // n+1
  std::vector<uint32_t> nb;
  double rlist2 = 0;
  bool list_valid = false;

  int64_t ncells() const { return (int64_t)nc[0] * nc[1] * nc[2]; }
  int threads() const { return nthreads > 0 ? nthreads : omp_get_max_threads(); }

  void pair_dr(int64_t i, uint32_t e, double d[3]) const
  {
    const uint32_t j = xsmd_nbr_index(e);
    d[0] = xsmd_image(x[j], xsmd_nbr_img_x(e), L[0]) - x[i];
    d[1] = xsmd_image(y[j], xsmd_nbr_img_y(e), L[1]) - y[i];
    d[2] = xsmd_image(z[j], xsmd_nbr_img_z(e), L[2]) - z[i];
    if (xf_on) xsmd_xform_apply(xf, &d[0], &d[1], &d[2]);
  }
};

namespace {

template <class F>
void for_each_cell(const or_ctx& c, F&& f)
{
  const int64_t nc = c.ncells();
#pragma omp parallel for schedule(dynamic) num_threads(c.threads())
  for (int64_t cell = 0; cell < nc; cell++)
    for (int64_t i = c.cell_start[cell]; i < c.cell_start[cell + 1]; i++) f(i);
}

template <class T>
void permute(std::vector<T>& v, const std::vector<int64_t>& order, size_t stride = 1)
{
  std::vector<T> tmp(v.size());
  for (size_t k = 0; k < order.size(); k++)
    for (size_t s = 0; s < stride; s++) tmp[k * stride + s] = v[(size_t)order[k] * stride + s];
  v.swap(tmp);
}

// The 27-cell stencil visit shared by list build and nothing else: dz outer,
// dx inner; a neighbour cell outside a periodic axis wraps with image -1/+1.
template <class F>
void stencil(const or_ctx& c, int cx, int cy, int cz, F&& f)
{
  for (int dz = -1; dz <= 1; dz++) {
    int z = cz + dz, kz = 0;
    if (z < 0) { if (!c.periodic[2]) continue; z += c.nc[2]; kz = -1; }
    else if (z >= c.nc[2]) { if (!c.periodic[2]) continue; z -= c.nc[2]; kz = 1; }
    for (int dy = -1; dy <= 1; dy++) {
      int y = cy + dy, ky = 0;
      if (y < 0) { if (!c.periodic[1]) continue; y += c.nc[1]; ky = -1; }
      else if (y >= c.nc[1]) { if (!c.periodic[1]) continue; y -= c.nc[1]; ky = 1; }
      for (int dx = -1; dx <= 1; dx++) {
        int x = cx + dx, kx = 0;
        if (x < 0) { if (!c.periodic[0]) continue; x += c.nc[0]; kx = -1; }
        else if (x >= c.nc[0]) { if (!c.periodic[0]) continue; x -= c.nc[0]; kx = 1; }
        const int64_t cell = ((int64_t)z * c.nc[1] + y) * c.nc[0] + x;
        f(cell, kx, ky, kz);
      }
    }
  }
}

}  // namespace

// ==========================================================================
// C API (C linkage comes from the declarations in oracle.h)
// ==========================================================================

double or_eV(void) { return kEV; }
double or_kB(void) { return kKB; }

int or_interpolate(int n, double delta, const double* f, double* coeffs)
{
  if (n < 5 || !(delta > 0) || !f || !coeffs) return -1;
  interpolate(n, delta, f, coeffs);
  return 0;
}

// setfl reader: YAML::convert<EamAlloyParameters>::decode, eam_alloy.cpp:66-201
int or_setfl_header(const char* path, int* nelements, int* nrho, double* drho, int* nr,
                    double* dr, double* rcut)
{
  std::ifstream file(path);
  if (!file) return -1;
  for (int i = 1; i < 4; ++i) file.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  size_t nel = 0;
  file >> nel;
  std::string name;
  for (size_t i = 0; i < nel; i++) file >> name;
  size_t a, b;
  double c, d, e;
  file >> a >> c >> b >> d >> e;
  if (!file) return -2;
  *nelements = (int)nel;
  *nrho = (int)a;
  *drho = c;
  *nr = (int)b;
  *dr = d;
  *rcut = e;
  return 0;
}

int or_setfl_data(const char* path, double* frho, double* rhor, double* z2r, double* mass)
{
  int nel, nrho, nr;
  double drho, dr, rc;
  if (or_setfl_header(path, &nel, &nrho, &drho, &nr, &dr, &rc)) return -1;
  std::ifstream file(path);
  for (int i = 1; i < 4; ++i) file.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  size_t tmp;
  std::string name;
  file >> tmp;
  for (int i = 0; i < nel; i++) file >> name;
  double dd;
  file >> tmp >> dd >> tmp >> dd >> dd;
  for (int i = 0; i < nel; i++) {
    int zmat;
    double massmat, azero;
    std::string structmat;
    file >> std::ws;
    file >> zmat >> massmat >> azero >> structmat;
    mass[i] = massmat;
    for (int k = 0; k < nrho; k++) file >> frho[(size_t)i * nrho + k];
    for (int k = 0; k < nr; k++) file >> rhor[(size_t)i * nr + k];
  }
  const int nz2r = nel * (nel + 1) / 2;
  for (int i = 0; i < nz2r; i++)
    for (int k = 0; k < nr; k++) file >> z2r[(size_t)i * nr + k];
  return file ? 0 : -2;
}

or_ctx* or_create(void) { return new or_ctx(); }
void or_destroy(or_ctx* c) { delete c; }
int or_set_threads(or_ctx* c, int nthreads)
{
  c->nthreads = nthreads;
  return 0;
}

int or_set_system(or_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                  const double* vx, const double* vy, const double* vz, const uint8_t* type,
                  const double lo[3], const double hi[3], const int periodic[3], int nspecies,
                  const double* masses)
{
  if (n < 0 || n >= XSMD_NBR_MAX_ATOMS || nspecies < 1 || nspecies > 7) return -1;
  c->n = n;
  c->x.assign(x, x + n);
  c->y.assign(y, y + n);
  c->z.assign(z, z + n);
  c->vx.assign(vx, vx + n);
  c->vy.assign(vy, vy + n);
  c->vz.assign(vz, vz + n);
  c->type.assign(type, type + n);
  c->id.resize(n);
  std::iota(c->id.begin(), c->id.end(), (int64_t)0);
  c->ax.assign(n, 0.0);
  c->ay.assign(n, 0.0);
  c->az.assign(n, 0.0);
  c->ep.assign(n, 0.0);
  c->vir.assign(9 * (size_t)n, 0.0);
  c->rho.assign(n, 0.0);
  c->xr = c->x;
  c->yr = c->y;
  c->zr = c->z;
  for (int a = 0; a < 3; a++) {
    c->lo[a] = lo[a];
    c->hi[a] = hi[a];
    c->L[a] = hi[a] - lo[a];
    c->periodic[a] = periodic[a];
  }
  c->mass.assign(masses, masses + nspecies);
  c->list_valid = false;
  // until the first build, a single cell holds every atom
  c->nc[0] = c->nc[1] = c->nc[2] = 1;
  c->cell_start = {0, n};
  return 0;
}

// LJ operator set-up: LennardJonesParms (lennard_jones.h:33-37) and the
// energy shift ecut = E(rcut) from energy_cutoff (pair_potential_impl.hxx:487-498).
int or_set_lj(or_ctx* c, double epsilon, double sigma, double rcut)
{
  c->has_lj = true;
  c->lj_eps = epsilon;
  c->lj_sigma = sigma;
  c->lj_rcut = rcut;
  double e = 0.0, de = 0.0;
  if (rcut > 0.0) lj_compute_energy(epsilon, sigma, rcut, e, de);
  c->lj_ecut = e;
  return 0;
}

// lj_multi_force set-up: the per-type-pair parameter table of
// pair_potential_impl.hxx:282-320 (each pair's energy_cutoff at its own rcut).
int or_set_lj_multi(or_ctx* c, int ntypes, const double* epsilon, const double* sigma, const double* rcut)
{
  if (ntypes < 1 || ntypes > 7) return -1;
  const int np = ntypes * (ntypes + 1) / 2;
  c->ljm_ntypes = ntypes;
  c->ljm_eps.assign(epsilon, epsilon + np);
  c->ljm_sigma.assign(sigma, sigma + np);
  c->ljm_rcut.assign(rcut, rcut + np);
  c->ljm_ecut.assign(np, 0.0);
  for (int k = 0; k < np; k++) {
    double e = 0.0, de = 0.0;
    if (rcut[k] > 0.0) lj_compute_energy(epsilon[k], sigma[k], rcut[k], e, de);
    c->ljm_ecut[k] = e;
  }
  return 0;
}

// Table set-up as the tail of eam_alloy.cpp:112-201 (rdr, rdrho, rhomax, splines)
int or_set_eam(or_ctx* c, int nel, int nrho, double drho, int nr, double dr, double rcut,
               const double* frho, const double* rhor, const double* z2r)
{
  if (nel < 1 || nel > 7 || nrho < 5 || nr < 5) return -1;
  Eam& e = c->eam;
  e = Eam();
  e.nel = nel;
  e.nr = nr;
  e.nrho = nrho;
  e.rhomax = (nrho - 1) * drho;
  e.rc = rcut;
  e.rdr = 1.0 / dr;
  e.rdrho = 1.0 / drho;
  const int nz2r = nel * (nel + 1) / 2;
  auto mk = [](int n, double delta, const double* f) {
    Spline s;
    s.n = n;
    s.c.resize(8 * (size_t)(n + 1));
    interpolate(n, delta, f, s.c.data());
    return s;
  };
  for (int i = 0; i < nel; i++) e.frho.push_back(mk(nrho, drho, frho + (size_t)i * nrho));
  for (int i = 0; i < nel; i++) e.rhor.push_back(mk(nr, dr, rhor + (size_t)i * nr));
  for (int i = 0; i < nz2r; i++) e.z2r.push_back(mk(nr, dr, z2r + (size_t)i * nr));
  e.init_types(nel);
  c->has_eam = true;
  return 0;
}

// Neighbour-list rebuild: move_particles (wrap into the periodic box), cell
// binning with a stable in-cell order, then the full (non-Newton) per-atom
// list of every (atom, image) within rcut + skin -- the set chunk_neighbors
// stores at chunk_size 1 (data/config/config_move_particles.msp:54-68,
// iterator semantics src/debug/atom_neighbors_stats.cpp:169-227,
// filter signature src/particle_species/type_pair_rcut_neighbors.cpp:49-62).
// johnson_force parameters (19 doubles, EamJohnsonParameters order, internal units)
int or_set_johnson(or_ctx* c, const double* params, double rcut)
{
  if (!params || !(rcut > 0) || !(params[J_RE] > 0) || !(params[J_RHOE] > 0)) return -1;
  std::copy(params, params + 19, c->jp);
  c->johnson_rc = rcut;
  c->has_johnson = true;
  return 0;
}

int or_build_neighbors(or_ctx* c, double rcut, double skin)
{
  const int64_t n = c->n;
  const double rl = rcut + skin;
  for (int a = 0; a < 3; a++) {
    c->nc[a] = xsmd_cells_along(c->L[a], c->xf_on ? rl * xsmd_xform_reach(c->xfi, a) : rl);
    c->inv_cs[a] = c->nc[a] / c->L[a];
  }
  for (int64_t i = 0; i < n; i++) {
    if (c->periodic[0]) c->x[i] = xsmd_wrap(c->x[i], c->lo[0], c->L[0]);
    if (c->periodic[1]) c->y[i] = xsmd_wrap(c->y[i], c->lo[1], c->L[1]);
    if (c->periodic[2]) c->z[i] = xsmd_wrap(c->z[i], c->lo[2], c->L[2]);
  }
  std::vector<int64_t> cell(n);
  for (int64_t i = 0; i < n; i++) {
    const int cx = xsmd_cell_coord(c->x[i], c->lo[0], c->inv_cs[0], c->nc[0]);
    const int cy = xsmd_cell_coord(c->y[i], c->lo[1], c->inv_cs[1], c->nc[1]);
    const int cz = xsmd_cell_coord(c->z[i], c->lo[2], c->inv_cs[2], c->nc[2]);
    cell[i] = ((int64_t)cz * c->nc[1] + cy) * c->nc[0] + cx;
  }
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), (int64_t)0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return cell[a] < cell[b]; });
  permute(c->x, order);
  permute(c->y, order);
  permute(c->z, order);
  permute(c->vx, order);
  permute(c->vy, order);
  permute(c->vz, order);
  permute(c->ax, order);
  permute(c->ay, order);
  permute(c->az, order);
  permute(c->ep, order);
  permute(c->rho, order);
  permute(c->vir, order, 9);
  permute(c->type, order);
  permute(c->id, order);
  permute(cell, order);

  const int64_t ncell = c->ncells();
  c->cell_start.assign(ncell + 1, 0);
  for (int64_t i = 0; i < n; i++) c->cell_start[cell[i] + 1]++;
  for (int64_t k = 0; k < ncell; k++) c->cell_start[k + 1] += c->cell_start[k];

  c->xr = c->x;
  c->yr = c->y;
  c->zr = c->z;

  c->rlist2 = rl * rl;
  const double rl2 = c->rlist2;
  std::vector<int64_t> cnt(n + 1, 0);
  auto visit = [&](int64_t i, auto&& emit) {
    const int cx = (int)(cell[i] % c->nc[0]);
    const int cy = (int)((cell[i] / c->nc[0]) % c->nc[1]);
    const int cz = (int)(cell[i] / ((int64_t)c->nc[0] * c->nc[1]));
    stencil(*c, cx, cy, cz, [&](int64_t cb, int kx, int ky, int kz) {
      for (int64_t j = c->cell_start[cb]; j < c->cell_start[cb + 1]; j++) {
        if (j == i && kx == 0 && ky == 0 && kz == 0) continue;
        double dx = xsmd_image(c->x[j], kx, c->L[0]) - c->x[i];
        double dy = xsmd_image(c->y[j], ky, c->L[1]) - c->y[i];
        double dz = xsmd_image(c->z[j], kz, c->L[2]) - c->z[i];
        if (c->xf_on) xsmd_xform_apply(c->xf, &dx, &dy, &dz);
        if (xsmd_d2(dx, dy, dz) < rl2) emit(xsmd_nbr_encode((uint32_t)j, kx, ky, kz));
      }
    });
  };
  for_each_cell(*c, [&](int64_t i) {
    int64_t k = 0;
    visit(i, [&](uint32_t) { k++; });
    cnt[i + 1] = k;
  });
  for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  c->nb_off = cnt;
  c->nb.assign(cnt[n], 0);
  for_each_cell(*c, [&](int64_t i) {
    int64_t k = c->nb_off[i];
    visit(i, [&](uint32_t e) { c->nb[k++] = e; });
  });
  c->list_valid = true;
  return 0;
}

// zero_force_energy, src/compute/zero_force_energy.cu:38-74
int or_zero_force_energy(or_ctx* c)
{
  std::fill(c->ax.begin(), c->ax.end(), 0.0);
  std::fill(c->ay.begin(), c->ay.end(), 0.0);
  std::fill(c->az.begin(), c->az.end(), 0.0);
  std::fill(c->ep.begin(), c->ep.end(), 0.0);
  std::fill(c->vir.begin(), c->vir.end(), 0.0);
  return 0;
}

// force_to_accel, src/compute/force_to_accel.cu:38-55
int or_force_to_accel(or_ctx* c)
{
  for_each_cell(*c, [&](int64_t i) {
    const double m = c->mass[c->type[i]];
    c->ax[i] /= m;
    c->ay[i] /= m;
    c->az[i] /= m;
  });
  return 0;
}

namespace {

// LJ force: ForceOp over the full list, pair_potential_force_op_singlemat.h:45-146
// with the per-atom accumulation of force_op_impl2.hxx:22-78.
void lj_force(or_ctx* c, int flags)
{
  const double rc2 = c->lj_rcut * c->lj_rcut;
  for_each_cell(*c, [&](int64_t i) {
    double _ep = 0., _fx = 0., _fy = 0., _fz = 0.;
    double _vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t k = c->nb_off[i]; k < c->nb_off[i + 1]; k++) {
      double dr[3];
      c->pair_dr(i, c->nb[k], dr);
      const double d2 = xsmd_d2(dr[0], dr[1], dr[2]);
      if (d2 > rc2) continue;
      const double r = std::sqrt(d2);
      double e = 0.0, de = 0.0;
      lj_compute_energy(c->lj_eps, c->lj_sigma, r, e, de);
      e -= c->lj_ecut;
      de /= r;
      const double fe[3] = {de * dr[0], de * dr[1], de * dr[2]};
      _fx += fe[0];
      _fy += fe[1];
      _fz += fe[2];
      _ep += .5 * e;
      add_tensor_scaled(_vir, fe, dr, -0.5);
    }
    c->ax[i] += _fx;
    c->ay[i] += _fy;
    c->az[i] += _fz;
    if (flags & OR_ENERGY) c->ep[i] += _ep;
    if (flags & OR_VIRIAL)
      for (int q = 0; q < 9; q++) c->vir[9 * (size_t)i + q] += _vir[q];
  });
}

// lj_multi_force: PairMultiForceOp (pair_potential_force_op_multiparam.h:75-113) over
// the full list: the pair's own parameters, r <= its rcut, e shifted by its ecut.
void lj_multi_force(or_ctx* c, int flags)
{
  for_each_cell(*c, [&](int64_t i) {
    double _ep = 0., _fx = 0., _fy = 0., _fz = 0.;
    double _vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int ta = c->type[i];
    for (int64_t k = c->nb_off[i]; k < c->nb_off[i + 1]; k++) {
      const int64_t j = xsmd_nbr_index(c->nb[k]);
      const int tb = c->type[j];
      const int lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
      const int pid = hi * (hi + 1) / 2 + lo;
      double dr[3];
      c->pair_dr(i, c->nb[k], dr);
      const double d2 = xsmd_d2(dr[0], dr[1], dr[2]);
      const double r = std::sqrt(d2);
      if (!(r <= c->ljm_rcut[pid])) continue;
      double e = 0.0, de = 0.0;
      lj_compute_energy(c->ljm_eps[pid], c->ljm_sigma[pid], r, e, de);
      e -= c->ljm_ecut[pid];
      de /= r;
      const double fe[3] = {de * dr[0], de * dr[1], de * dr[2]};
      _fx += fe[0];
      _fy += fe[1];
      _fz += fe[2];
      _ep += .5 * e;
      add_tensor_scaled(_vir, fe, dr, -0.5);
    }
    c->ax[i] += _fx;
    c->ay[i] += _fy;
    c->az[i] += _fz;
    if (flags & OR_ENERGY) c->ep[i] += _ep;
    if (flags & OR_VIRIAL)
      for (int q = 0; q < 9; q++) c->vir[9 * (size_t)i + q] += _vir[q];
  });
}

// ---- Johnson EAM (johnson.h:59-167)

// c exp(-a (x - 1)) / (1 + (x - k)^20) at x = r / re and its r-derivative
void johnson_screened(double c, double a, double k, double x, double ire, double& v, double& dv)
{
  const double num = c * std::exp(-a * (x - 1.0));
  const double s = x - k;
  const double s20 = std::pow(s, 20);
  const double den = 1.0 + s20;
  v = num / den;
  dv = ire * ((-a * num) * den - num * (20 * s20 / s)) / (den * den);
}

// eam_johnson_rho (johnson.h:105-124)
void johnson_rho(const double* p, double r, double& rho, double& drho)
{
  johnson_screened(p[J_FE], p[J_BETA], p[J_LAMBDA], r / p[J_RE], 1.0 / p[J_RE], rho, drho);
}

// eam_johnson_phi (johnson.h:59-99): A-term + (-B)-term
void johnson_phi(const double* p, double r, double& phi, double& dphi)
{
  const double x = r / p[J_RE], ire = 1.0 / p[J_RE];
  double vb, db;
  johnson_screened(p[J_A], p[J_ALPHA], p[J_KAPPA], x, ire, phi, dphi);
  johnson_screened(-p[J_B], p[J_BETA], p[J_LAMBDA], x, ire, vb, db);
  phi += vb;
  dphi += db;
}

// eam_johnson_fEmbed (johnson.h:131-167)
void johnson_embed(const double* p, double rho, double& f, double& df)
{
  const double rhoe = p[J_RHOE], rhon = 0.85 * rhoe, rho0 = 1.15 * rhoe;
  if (rho < rhon) {
    const double t = rho / rhon - 1.0;
    f = p[J_FN0] + p[J_FN1] * t + p[J_FN2] * (t * t) + p[J_FN3] * (t * (t * t));
    df = (p[J_FN1] + 2.0 * p[J_FN2] * t + 3.0 * p[J_FN3] * (t * t)) * (1.0 / rhon);
  } else if (rho < rho0) {
    const double t = rho * (1.0 / rhoe) - 1.0;
    f = p[J_F0] + p[J_F1] * t + p[J_F2] * (t * t) + p[J_F3] * (t * (t * t));
    df = (p[J_F1] + 2.0 * p[J_F2] * t + 3.0 * p[J_F3] * (t * t)) * (1.0 / rhoe);
  } else {
    const double q = rho * (1.0 / rhoe);
    const double qe = std::pow(q, p[J_ETA]);
    const double lq = std::log(qe);
    f = p[J_FO] * (1.0 - lq) * qe;
    df = (-p[J_ETA] * qe + (1.0 - lq) * p[J_ETA] * qe) * (p[J_FO] * (1.0 / rhoe) / q);
  }
}

// johnson_force: EmbOp then ForceOp of the single-material template
// (eam_potential.cu:126-172, eam_force_op_singlemat.h:39-171).  Energy always
// accumulates (both ops add to ep unconditionally); the virial with OR_VIRIAL.
void johnson_force(or_ctx* c, int flags)
{
  const double* p = c->jp;
  const double rc2 = c->johnson_rc * c->johnson_rc;
  for_each_cell(*c, [&](int64_t i) {
    double rho = 0.0;
    for (int64_t k = c->nb_off[i]; k < c->nb_off[i + 1]; k++) {
      double dr[3];
      c->pair_dr(i, c->nb[k], dr);
      const double d2 = xsmd_d2(dr[0], dr[1], dr[2]);
      if (d2 > rc2) continue;
      double v, dv;
      johnson_rho(p, std::sqrt(d2), v, dv);
      rho += v;
    }
    double emb, demb;
    johnson_embed(p, rho, emb, demb);
    c->ep[i] += emb;
    c->rho[i] = demb;
  });
  for_each_cell(*c, [&](int64_t i) {
    double f[3] = {0., 0., 0.}, ep = 0.0;
    double vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t k = c->nb_off[i]; k < c->nb_off[i + 1]; k++) {
      double dr[3];
      c->pair_dr(i, c->nb[k], dr);
      const double d2 = xsmd_d2(dr[0], dr[1], dr[2]);
      if (d2 > rc2) continue;
      const double r = std::sqrt(d2);
      double rho, drho, phi, dphi;
      johnson_rho(p, r, rho, drho);
      johnson_phi(p, r, phi, dphi);
      const double de = (drho * (c->rho[i] + c->rho[xsmd_nbr_index(c->nb[k])]) + dphi) / r;
      const double fe[3] = {de * dr[0], de * dr[1], de * dr[2]};
      f[0] += fe[0];
      f[1] += fe[1];
      f[2] += fe[2];
      ep += .5 * phi;
      add_tensor_scaled(vir, fe, dr, -0.5);
    }
    c->ax[i] += f[0];
    c->ay[i] += f[1];
    c->az[i] += f[2];
    c->ep[i] += ep;
    if (flags & OR_VIRIAL)
      for (int q = 0; q < 9; q++) c->vir[9 * (size_t)i + q] += vir[q];
  });
}

// EAM pass 1: parallel_memset + SymRhoOp<false>, eam_potential_multimat.cu:186-203,
// eam_force_op_multimat.h:114-172
void eam_rho_pass(or_ctx* c)
{
  const double rc2 = c->eam.rc * c->eam.rc;
  std::fill(c->rho.begin(), c->rho.end(), 0.0);
  for_each_cell(*c, [&](int64_t i) {
    double rho = 0.0;
    const int ta = c->type[i];
    if (!c->eam_on(ta)) return;  // a disabled species keeps rho = 0 (eam_force_op_multimat.h:149)
    for (int64_t k = c->nb_off[i]; k < c->nb_off[i + 1]; k++) {
      double dr[3];
      c->pair_dr(i, c->nb[k], dr);
      const double d2 = xsmd_d2(dr[0], dr[1], dr[2]);
      if (d2 > rc2) continue;
      const int tb = c->type[xsmd_nbr_index(c->nb[k])];
      if (!c->eam_on(tb)) continue;  // pair disabled (:151-153)
      const double r = std::sqrt(d2);
      rho += eam_rho_noderiv(c->eam, r, tb, ta);
    }
    c->rho[i] += rho;
  });
}

// EAM pass 1b: Rho2EmbOp, eam_force_op_multimat.h:174-197
void eam_rho2emb_pass(or_ctx* c, bool energy)
{
  for_each_cell(*c, [&](int64_t i) {
    const int ta = c->type[i];
    if (!c->eam_on(ta)) return;  // Rho2EmbOp skips a disabled species (eam_force_op_multimat.h:180,191)
    if (energy) {
      double emb = 0., demb = 0.;
      eam_fembed(c->eam, c->rho[i], emb, demb, ta);
      c->rho[i] = demb;
      c->ep[i] += emb;
    } else {
      c->rho[i] = eam_fembed_deriv(c->eam, c->rho[i], ta);
    }
  });
}

// EAM pass 2: SymForceOp<false>, eam_force_op_multimat.h:213-343
void eam_force_pass(or_ctx* c, int flags)
{
  const double rc2 = c->eam.rc * c->eam.rc;
  const bool eflag = flags & (OR_ENERGY | OR_VIRIAL);
  for_each_cell(*c, [&](int64_t i) {
    const double fpi = c->rho[i];
    const int ta = c->type[i];
    double f[3] = {0., 0., 0.};
    double ep = 0.0;
    double vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t kend = c->eam_on(ta) ? c->nb_off[i + 1] : c->nb_off[i];  // (eam_force_op_multimat.h:303)
    for (int64_t k = c->nb_off[i]; k < kend; k++) {
      double dr[3];
      c->pair_dr(i, c->nb[k], dr);
      const double d2 = xsmd_d2(dr[0], dr[1], dr[2]);
      if (d2 > rc2) continue;
      const uint32_t j = xsmd_nbr_index(c->nb[k]);
      const int tb = c->type[j];
      if (!c->eam_on(tb)) continue;  // (:308)
      const double r = std::sqrt(d2);
      double phi = 0.;
      double dr_fe[3] = {dr[0], dr[1], dr[2]};
      eam_mm_force(c->eam, dr_fe, phi, r, fpi, c->rho[j], ta, tb);
      f[0] += dr_fe[0];
      f[1] += dr_fe[1];
      f[2] += dr_fe[2];
      if (eflag) {
        ep += .5 * phi;
        add_tensor_scaled(vir, dr_fe, dr, -0.5);
      }
    }
    c->ax[i] += f[0];
    c->ay[i] += f[1];
    c->az[i] += f[2];
    if (flags & OR_ENERGY) c->ep[i] += ep;
    if (flags & OR_VIRIAL)
      for (int q = 0; q < 9; q++) c->vir[9 * (size_t)i + q] += vir[q];
  });
}

}  // namespace

int or_set_eam_types(or_ctx* c, unsigned mask)
{
  c->eam_types = mask;
  return 0;
}

int or_eam_pass(or_ctx* c, int eam_rho, int eam_rho2emb, int eam_force, int flags)
{
  if (!c->has_eam || !c->list_valid) return -1;
  if (eam_rho) eam_rho_pass(c);
  if (eam_rho2emb) eam_rho2emb_pass(c, flags & OR_ENERGY);
  if (eam_force) eam_force_pass(c, flags);
  return 0;
}

// compute_all_forces_energy: zero_force_energy, compute_force, force_to_accel
// (data/config/config_numerical_schemes.msp:8-15)
int or_compute_forces(or_ctx* c, int potential, int flags)
{
  if (!c->list_valid) return -1;
  or_zero_force_energy(c);
  if (potential == OR_POT_LJ) {
    if (!c->has_lj) return -2;
    lj_force(c, flags);
  } else if (potential == OR_POT_LJ_MULTI) {
    if (c->ljm_ntypes < 1) return -2;
    lj_multi_force(c, flags);
  } else if (potential == OR_POT_JOHNSON) {
    if (!c->has_johnson) return -2;
    johnson_force(c, flags);
  } else {
    if (!c->has_eam) return -2;
    eam_rho_pass(c);
    eam_rho2emb_pass(c, flags & OR_ENERGY);
    eam_force_pass(c, flags);
  }
  or_force_to_accel(c);
  return 0;
}

// push_f_v_r: x += v*dt + a*dt^2/2 (dt already multiplied by dt_scale); with
// xform_mode INV_XFORM and a non-identity xform the step is mapped back to grid
// space first, x += inv_xform * (v*dt + a*dt^2/2)
// (src/numerical_schemes/push_vec3_2nd_order_xform_minimization.cpp:34).
int or_push_f_v_r(or_ctx* c, double dt)
{
  return or_push_f_v_r_mode(c, dt, 1);
}

int or_push_f_v_r_mode(or_ctx* c, double dt, int inv_xform)
{
  const double dt2 = dt * dt * 0.5;
  const bool map = inv_xform && c->xf_on;
  for_each_cell(*c, [&](int64_t i) {
    double sx = std::fma(c->vx[i], dt, c->ax[i] * dt2);
    double sy = std::fma(c->vy[i], dt, c->ay[i] * dt2);
    double sz = std::fma(c->vz[i], dt, c->az[i] * dt2);
    if (map) xsmd_xform_apply(c->xfi, &sx, &sy, &sz);
    c->x[i] += sx;
    c->y[i] += sy;
    c->z[i] += sz;
  });
  return 0;
}

// Domain xform: m (row-major) maps grid to physical coordinates; NULL or the
// identity selects the plain traversal.  Returns -1 for a singular or reflecting m.
int or_set_xform(or_ctx* c, const double* m)
{
  static const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (!m) m = id;
  const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                     m[2] * (m[3] * m[7] - m[4] * m[6]);
  if (!(det > 0.0)) return -1;
  const double inv[9] = {(m[4] * m[8] - m[5] * m[7]) / det, (m[2] * m[7] - m[1] * m[8]) / det,
                         (m[1] * m[5] - m[2] * m[4]) / det, (m[5] * m[6] - m[3] * m[8]) / det,
                         (m[0] * m[8] - m[2] * m[6]) / det, (m[2] * m[3] - m[0] * m[5]) / det,
                         (m[3] * m[7] - m[4] * m[6]) / det, (m[1] * m[6] - m[0] * m[7]) / det,
                         (m[0] * m[4] - m[1] * m[3]) / det};
  bool ident = true;
  for (int q = 0; q < 9; q++) ident = ident && m[q] == id[q];
  c->xf_on = !ident;
  std::copy(m, m + 9, c->xf);
  std::copy(inv, inv + 9, c->xfi);
  c->list_valid = false;
  return 0;
}

// push_f_v: v += a*dt (docs/notes/markdown/md.md:64-75)
int or_push_f_v(or_ctx* c, double dt)
{
  for_each_cell(*c, [&](int64_t i) {
    c->vx[i] = std::fma(c->ax[i], dt, c->vx[i]);
    c->vy[i] = std::fma(c->ay[i], dt, c->vy[i]);
    c->vz[i] = std::fma(c->az[i], dt, c->vz[i]);
  });
  return 0;
}

// particle_displ_over: any displacement since the last rebuild above the
// threshold (config_move_particles.msp:19-23,116-118).
int or_displ_over(or_ctx* c, double threshold)
{
  const double t2 = threshold * threshold;
  int over = 0;
#pragma omp parallel for reduction(| : over) num_threads(c->threads())
  for (int64_t i = 0; i < c->n; i++) {
    double dx = c->x[i] - c->xr[i], dy = c->y[i] - c->yr[i], dz = c->z[i] - c->zr[i];
    if (c->xf_on) xsmd_xform_apply(c->xf, &dx, &dy, &dz);  // the threshold is a physical distance
    if (xsmd_d2(dx, dy, dz) > t2) over = 1;
  }
  return over;
}

// simulation_thermodynamic_state reduction, simulation_thermodynamic_state.cpp:101-163:
// out = virial(9), ke_tensor(9), momentum(3), kinetic_energy(3), potential_energy,
// mass, particle count.  Per-cell partial sums, then summed in cell order.
int or_thermo(or_ctx* c, double out[27])
{
  const int64_t nc = c->ncells();
  std::vector<double> part((size_t)nc * 26, 0.0);
#pragma omp parallel for schedule(dynamic) num_threads(c->threads())
  for (int64_t cell = 0; cell < nc; cell++) {
    double* p = &part[(size_t)cell * 26];
    for (int64_t i = c->cell_start[cell]; i < c->cell_start[cell + 1]; i++) {
      const double m = c->mass[c->type[i]];
      const double v[3] = {c->vx[i], c->vy[i], c->vz[i]};
      for (int q = 0; q < 9; q++) p[q] += c->vir[9 * (size_t)i + q];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) p[9 + 3 * a + b] += (v[a] * v[b]) * m;
      for (int a = 0; a < 3; a++) p[18 + a] += v[a] * m;
      for (int a = 0; a < 3; a++) p[21 + a] += v[a] * v[a] * m;
      p[24] += c->ep[i];
      p[25] += m;
    }
  }
  double s[26] = {0};
  for (int64_t cell = 0; cell < nc; cell++)
    for (int q = 0; q < 26; q++) s[q] += part[(size_t)cell * 26 + q];
  for (int q = 0; q < 9; q++) out[q] = s[q];
  for (int q = 9; q < 18; q++) out[q] = 0.5 * s[q];
  for (int q = 18; q < 21; q++) out[q] = s[q];
  for (int q = 21; q < 24; q++) out[q] = 0.5 * s[q];
  out[24] = s[24];
  out[25] = s[25];
  out[26] = (double)c->n;
  return 0;
}

// NVE loop: first_iteration then verlet_nve steps
// (data/config/main-config.msp:145-170; config_numerical_schemes.msp:23-28,44-52;
//  config_move_particles.msp:116-152 with max_displ = rcut_inc / 2).
int or_md_run(or_ctx* c, int potential, int nsteps, double dt, double rcut, double skin,
              double* thermo_out)
{
  const int eflags = thermo_out ? (OR_ENERGY | OR_VIRIAL) : 0;
  or_build_neighbors(c, rcut, skin);
  int rc = or_compute_forces(c, potential, eflags);
  if (rc) return rc;
  if (thermo_out) or_thermo(c, thermo_out);
  rc = or_md_steps(c, potential, nsteps, dt, rcut, skin, thermo_out ? thermo_out + 27 : nullptr);
  return rc < 0 ? rc : rc + 1;
}

// The verlet_nve steps alone, continuing from current accelerations and list;
// returns the number of rebuilds.  thermo_out (may be NULL) receives nsteps x 27.
int or_md_steps(or_ctx* c, int potential, int nsteps, double dt, double rcut, double skin,
                double* thermo_out)
{
  if (!c->list_valid) return -1;
  const int eflags = thermo_out ? (OR_ENERGY | OR_VIRIAL) : 0;
  int nrebuild = 0;
  for (int step = 1; step <= nsteps; step++) {
    or_push_f_v_r(c, dt * 1.0);
    or_push_f_v(c, dt * 0.5);
    if (or_displ_over(c, 0.5 * skin)) {
      or_build_neighbors(c, rcut, skin);
      nrebuild++;
    }
    or_compute_forces(c, potential, eflags);
    or_push_f_v(c, dt * 0.5);
    if (thermo_out) or_thermo(c, thermo_out + 27 * (size_t)(step - 1));
  }
  return nrebuild;
}

int64_t or_n(const or_ctx* c) { return c->n; }

int or_get_state(const or_ctx* c, double* x, double* y, double* z, double* vx, double* vy,
                 double* vz, double* ax, double* ay, double* az, double* ep, double* vir9,
                 double* rho_dEmb)
{
  for (int64_t k = 0; k < c->n; k++) {
    const int64_t i = c->id[k];
    if (x) x[i] = c->x[k];
    if (y) y[i] = c->y[k];
    if (z) z[i] = c->z[k];
    if (vx) vx[i] = c->vx[k];
    if (vy) vy[i] = c->vy[k];
    if (vz) vz[i] = c->vz[k];
    if (ax) ax[i] = c->ax[k];
    if (ay) ay[i] = c->ay[k];
    if (az) az[i] = c->az[k];
    if (ep) ep[i] = c->ep[k];
    if (vir9)
      for (int q = 0; q < 9; q++) vir9[9 * i + q] = c->vir[9 * (size_t)k + q];
    if (rho_dEmb) rho_dEmb[i] = c->rho[k];
  }
  return 0;
}

int or_get_order(const or_ctx* c, int64_t* perm)
{
  for (int64_t k = 0; k < c->n; k++) perm[k] = c->id[k];
  return 0;
}

int64_t or_neighbor_count(const or_ctx* c, int64_t* counts)
{
  if (!c->list_valid) return -1;
  for (int64_t k = 0; k < c->n; k++) counts[c->id[k]] = c->nb_off[k + 1] - c->nb_off[k];
  return c->nb_off[c->n];
}

int or_neighbor_get(const or_ctx* c, int64_t* ids, int32_t* imgs)
{
  if (!c->list_valid) return -1;
  std::vector<int64_t> off(c->n + 1, 0);
  for (int64_t k = 0; k < c->n; k++) off[c->id[k] + 1] = c->nb_off[k + 1] - c->nb_off[k];
  for (int64_t i = 0; i < c->n; i++) off[i + 1] += off[i];
  for (int64_t k = 0; k < c->n; k++) {
    int64_t o = off[c->id[k]];
    for (int64_t q = c->nb_off[k]; q < c->nb_off[k + 1]; q++, o++) {
      const uint32_t e = c->nb[q];
      ids[o] = c->id[xsmd_nbr_index(e)];
      if (imgs) {
        imgs[3 * o + 0] = xsmd_nbr_img_x(e);
        imgs[3 * o + 1] = xsmd_nbr_img_y(e);
        imgs[3 * o + 2] = xsmd_nbr_img_z(e);
      }
    }
  }
  return 0;
}

namespace {
template <class F>
void brute(const or_ctx* c, double rlist, F&& f)
{
  const double rl2 = rlist * rlist;
  for (int64_t i = 0; i < c->n; i++)
    for (int64_t j = 0; j < c->n; j++)
      for (int kz = -c->periodic[2]; kz <= c->periodic[2]; kz++)
        for (int ky = -c->periodic[1]; ky <= c->periodic[1]; ky++)
          for (int kx = -c->periodic[0]; kx <= c->periodic[0]; kx++) {
            if (i == j && kx == 0 && ky == 0 && kz == 0) continue;
            double dx = xsmd_image(c->x[j], kx, c->L[0]) - c->x[i];
            double dy = xsmd_image(c->y[j], ky, c->L[1]) - c->y[i];
            double dz = xsmd_image(c->z[j], kz, c->L[2]) - c->z[i];
            if (c->xf_on) xsmd_xform_apply(c->xf, &dx, &dy, &dz);
            if (xsmd_d2(dx, dy, dz) < rl2) f(i, j, kx, ky, kz);
          }
}
}  // namespace

int64_t or_brute_neighbor_count(const or_ctx* c, double rlist, int64_t* counts)
{
  for (int64_t i = 0; i < c->n; i++) counts[i] = 0;
  int64_t tot = 0;
  brute(c, rlist, [&](int64_t i, int64_t, int, int, int) {
    counts[c->id[i]]++;
    tot++;
  });
  return tot;
}

int or_brute_neighbor_get(const or_ctx* c, double rlist, int64_t* ids, int32_t* imgs)
{
  struct E { int64_t j; int kx, ky, kz; };
  std::vector<std::vector<E>> per(c->n);
  brute(c, rlist, [&](int64_t i, int64_t j, int kx, int ky, int kz) {
    per[c->id[i]].push_back({c->id[j], kx, ky, kz});
  });
  int64_t o = 0;
  for (int64_t i = 0; i < c->n; i++) {
    auto& v = per[i];
    std::sort(v.begin(), v.end(), [](const E& a, const E& b) {
      if (a.j != b.j) return a.j < b.j;
      if (a.kz != b.kz) return a.kz < b.kz;
      if (a.ky != b.ky) return a.ky < b.ky;
      return a.kx < b.kx;
    });
    for (auto& e : v) {
      ids[o] = e.j;
      if (imgs) {
        imgs[3 * o] = e.kx;
        imgs[3 * o + 1] = e.ky;
        imgs[3 * o + 2] = e.kz;
      }
      o++;
    }
  }
  return 0;
}

double or_eam_rho(const or_ctx* c, double r, int itype, int jtype)
{
  return eam_rho_noderiv(c->eam, r, itype, jtype);
}

void or_eam_pair(const or_ctx* c, double r, double fpi, double fpj, int itype, int jtype,
                 double* fpair_over_r, double* phi)
{
  double d[3] = {1.0, 0.0, 0.0};
  eam_mm_force(c->eam, d, *phi, r, fpi, fpj, itype, jtype);
  *fpair_over_r = d[0];
}

void or_eam_embed(const or_ctx* c, double rho, int itype, double* F, double* Fp)
{
  eam_fembed(c->eam, rho, *F, *Fp, itype);
}

void or_lj_energy(const or_ctx* c, double r, double* e, double* de)
{
  lj_compute_energy(c->lj_eps, c->lj_sigma, r, *e, *de);
}

void or_johnson_eval(const double* params, double r, double out[4])
{
  johnson_rho(params, r, out[0], out[1]);
  johnson_phi(params, r, out[2], out[3]);
}

void or_johnson_embed(const double* params, double rho, double* F, double* Fp)
{
  johnson_embed(params, rho, *F, *Fp);
}
