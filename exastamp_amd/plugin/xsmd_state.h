// xsmd_state.h -- shared state of the exaStamp operator plugin over libxsmd.
//
// The plugin is the reference-side binding of libxsmd: onika OperatorNode classes
// with the slots of the exaStamp/exaNBody operators they stand in for, each calling
// the C ABI of include/xsmd.h.  It is built inside an exaStamp tree (onika, exaNBody
// and yaml-cpp present) with onika_add_plugin; this repository cannot compile it
// (none of those is available here), so it is kept to the onika API exaStamp's own
// operators use (eam_potential_multimat.cu:20-31, pair_potential_impl.hxx:100-121,
// force_to_accel.cu:78-110).
//
// Ownership of the particle state.  One xs_ctx per process (per GPU) holds the atoms
// in its own cell order.  xsmd_chunk_neighbors copies the grid's non-ghost particles
// (positions, velocities, accelerations, types, in cell order) into it and builds the
// list; from then on the state lives on the GPU and the xsmd operators of the
// integration loop (push_f_v_r, push_f_v, zero_force_energy, the force operators,
// force_to_accel, simulation_thermodynamic_state) work there.  xsmd_sync_grid copies
// positions, velocities, accelerations (and ep / virial when allocated) back into
// the grid for any operator that reads the grid itself (particle moves, dumps).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include <onika/log.h>

#include "xsmd.h"

namespace exaStamp
{
  struct XsmdState
  {
    xs_ctx* ctx = nullptr;
    bool on_gpu = false;        // the ctx holds the current particle state
    bool tables_loaded = false; // eam tables uploaded from the current parameters
    int64_t n = 0;
    // where atom k of the ctx (its id) lives in the grid
    std::vector<uint32_t> cell_of;
    std::vector<uint32_t> index_in_cell;

    static XsmdState& instance()
    {
      static XsmdState s;
      return s;
    }
    xs_ctx* context()
    {
      if( ctx == nullptr ) check( xs_ctx_create( 0, &ctx ), "xs_ctx_create" );
      return ctx;
    }
    void check( xs_status st, const char* what ) const
    {
      if( st != XS_OK ) onika::fatal_error() << "xsmd " << what << ": "
                                             << ( ctx ? xs_last_error(ctx) : "no context" ) << std::endl;
    }
  };

  // Copies the grid's non-ghost particles into the ctx (cell order; id = rank in that
  // order) with the domain's box and periodicity.
  template<class GridT, class DomainT>
  inline void xsmd_upload( XsmdState& st, GridT& grid, const DomainT& domain )
  {
    auto cells = grid.cells();
    const size_t n_cells = grid.number_of_cells();
    std::vector<double> rx, ry, rz, vx, vy, vz;
    std::vector<uint8_t> type;
    st.cell_of.clear();
    st.index_in_cell.clear();
    for( size_t c = 0; c < n_cells; c++ )
    {
      if( grid.is_ghost_cell(c) ) continue;
      const size_t np = cells[c].size();
      for( size_t p = 0; p < np; p++ )
      {
        rx.push_back( cells[c][field::rx][p] ); ry.push_back( cells[c][field::ry][p] ); rz.push_back( cells[c][field::rz][p] );
        vx.push_back( cells[c][field::vx][p] ); vy.push_back( cells[c][field::vy][p] ); vz.push_back( cells[c][field::vz][p] );
        type.push_back( static_cast<uint8_t>( cells[c][field::type][p] ) );
        st.cell_of.push_back( c );
        st.index_in_cell.push_back( p );
      }
    }
    st.n = static_cast<int64_t>( rx.size() );
    const auto b = domain.bounds();
    const double lo[3] = { b.bmin.x, b.bmin.y, b.bmin.z };
    const double hi[3] = { b.bmax.x, b.bmax.y, b.bmax.z };
    const int periodic[3] = { domain.periodic_boundary_x(), domain.periodic_boundary_y(), domain.periodic_boundary_z() };
    st.check( xs_system_set( st.context(), st.n, rx.data(), ry.data(), rz.data(), vx.data(), vy.data(), vz.data(),
                             type.data(), nullptr, 1, lo, hi, periodic ), "xs_system_set" );
    // grid coordinates in, physical displacements through the domain's xform
    const auto X = domain.xform();
    const double m[9] = { X.m11, X.m12, X.m13, X.m21, X.m22, X.m23, X.m31, X.m32, X.m33 };
    st.check( xs_domain_xform( st.context(), domain.xform_is_identity() ? nullptr : m ), "xs_domain_xform" );
    st.on_gpu = true;
  }

  // Copies positions, velocities and accelerations (the fx..fz fields hold
  // accelerations after force_to_accel) back into the grid, and ep / virial when the
  // grid has them.
  template<class GridT>
  inline void xsmd_download( XsmdState& st, GridT& grid, bool with_ep, bool with_virial )
  {
    const int64_t n = st.n;
    std::vector<double> r[3], v[3], a(3 * n), ep, vir;
    for( int d = 0; d < 3; d++ ) { r[d].resize(n); v[d].resize(n); }
    st.check( xs_system_get( st.ctx, r[0].data(), r[1].data(), r[2].data(), v[0].data(), v[1].data(), v[2].data() ),
              "xs_system_get" );
    st.check( xs_field_get( st.ctx, XS_FIELD_ACCEL, a.data() ), "xs_field_get" );
    if( with_ep ) { ep.resize(n); st.check( xs_field_get( st.ctx, XS_FIELD_EP, ep.data() ), "xs_field_get" ); }
    if( with_virial ) { vir.resize(9 * n); st.check( xs_field_get( st.ctx, XS_FIELD_VIRIAL, vir.data() ), "xs_field_get" ); }
    auto cells = grid.cells();
    for( int64_t k = 0; k < n; k++ )
    {
      const size_t c = st.cell_of[k], p = st.index_in_cell[k];
      cells[c][field::rx][p] = r[0][k]; cells[c][field::ry][p] = r[1][k]; cells[c][field::rz][p] = r[2][k];
      cells[c][field::vx][p] = v[0][k]; cells[c][field::vy][p] = v[1][k]; cells[c][field::vz][p] = v[2][k];
      cells[c][field::fx][p] = a[k]; cells[c][field::fy][p] = a[n + k]; cells[c][field::fz][p] = a[2 * n + k];
      if( with_ep ) cells[c][field::ep][p] = ep[k];
      if( with_virial )
      {
        const double* m = &vir[9 * k];
        cells[c][field::virial][p] = onika::math::Mat3d{ m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8] };
      }
    }
  }
}
