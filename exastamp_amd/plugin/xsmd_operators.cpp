// xsmd_operators.cpp -- onika operators of the exaStamp short-range force path over
// libxsmd (include/xsmd.h).  Each class keeps the slots of the operator it stands in
// for; the factory names carry an xsmd_ prefix so that a deck selects them by
// rebinding the reference's names (for instance `compute_force: xsmd_eam_alloy_force`,
// `push_f_v_r: xsmd_push_f_v_r`), next to exaNBody's own registrations.
//
// Built in an exaStamp tree by the CMakeLists.txt beside this file (onika_add_plugin,
// linked to libxsmd.so).  Not compiled in this repository: onika, exaNBody and
// yaml-cpp are not available here.
#include <onika/scg/operator.h>
#include <onika/scg/operator_factory.h>
#include <onika/scg/operator_slot.h>
#include <onika/log.h>
#include <exanb/core/make_grid_variant_operator.h>
#include <exanb/core/domain.h>
#include <exanb/core/grid.h>
#include <exanb/particle_neighbors/chunk_neighbors.h>
#include <exaStamp/particle_species/particle_specie.h>
#include <exaStamp/potential/eam/eam_alloy.h>
#include <exaStamp/thermo_state/thermodynamic_state.h>

#include <algorithm>
#include <string>
#include <vector>

#include "xsmd_state.h"

namespace exaStamp
{
  using namespace exanb;

  // species masses into the ctx (particle_specie.h:52-122)
  inline void xsmd_species( XsmdState& st, const ParticleSpecies& species )
  {
    std::vector<double> m( species.size() );
    for( size_t s = 0; s < species.size(); s++ ) m[s] = species.at(s).m_mass;
    st.check( xs_species_set( st.context(), static_cast<int>( m.size() ), m.data() ), "xs_species_set" );
  }

  // chunk_neighbors (config_move_particles.msp:54-68): the grid's particles into the
  // ctx, then the library's cell-block list at rcut_max + rcut_inc.
  template<class GridT>
  class XsmdChunkNeighbors : public OperatorNode
  {
    ADD_SLOT( GridT           , grid     , INPUT_OUTPUT );
    ADD_SLOT( Domain          , domain   , INPUT , REQUIRED );
    ADD_SLOT( ParticleSpecies , species  , INPUT , REQUIRED );
    ADD_SLOT( double          , rcut_max , INPUT , 0.0 );
    ADD_SLOT( double          , rcut_inc , INPUT , 0.0 );
    ADD_SLOT( long            , chunk_size , INPUT , 1 );
  public:
    inline void execute() override final
    {
      if( ! domain->xform_is_identity() ) fatal_error() << "xsmd: only the identity domain transform is supported" << std::endl;
      XsmdState& st = XsmdState::instance();
      xsmd_species( st, *species );
      xsmd_upload( st, *grid, *domain );
      st.check( xs_chunk_neighbors( st.ctx, *rcut_max, *rcut_inc, static_cast<int>( *chunk_size ) ), "xs_chunk_neighbors" );
    }
  };

  // particle_displ_over (config_move_particles.msp:19-23) on the ctx's positions
  class XsmdParticleDisplOver : public OperatorNode
  {
    ADD_SLOT( double , threshold , INPUT , 0.0 );
    ADD_SLOT( bool   , result    , OUTPUT );
  public:
    inline void execute() override final
    {
      XsmdState& st = XsmdState::instance();
      int over = 1;
      if( st.on_gpu ) st.check( xs_particle_displ_over( st.ctx, *threshold, &over ), "xs_particle_displ_over" );
      *result = ( over != 0 );
    }
  };

  // eam_alloy_force (eam_potential_multimat.cu:65-259): slots of :88-109
  template<class GridT>
  class XsmdEamAlloyForce : public OperatorNode
  {
    using StringVector = std::vector<std::string>;
    ADD_SLOT( ParticleSpecies    , species          , INPUT , REQUIRED );
    ADD_SLOT( EamAlloyParameters , parameters       , INPUT_OUTPUT , REQUIRED );
    ADD_SLOT( StringVector       , types            , INPUT , StringVector{} );
    ADD_SLOT( double             , rcut             , INPUT );
    ADD_SLOT( double             , rcut_max         , INPUT_OUTPUT , 0.0 );
    ADD_SLOT( double             , ghost_dist_max   , INPUT_OUTPUT , 0.0 );
    ADD_SLOT( GridT              , grid             , INPUT_OUTPUT );
    ADD_SLOT( bool , trigger_thermo_state , INPUT , OPTIONAL );
    ADD_SLOT( bool , compute_virial , INPUT , false );
    ADD_SLOT( bool , eam_rho      , INPUT , true );
    ADD_SLOT( bool , eam_rho2emb  , INPUT , true );
    ADD_SLOT( bool , eam_ghost    , INPUT , true );
    ADD_SLOT( bool , eam_force    , INPUT , true );
    ADD_SLOT( bool , eam_symmetry , INPUT , false );
    ADD_SLOT( bool , mixed_precision , INPUT , true );
  public:
    inline void execute() override final
    {
      *rcut_max = std::max( *rcut , *rcut_max );
      if( ( *eam_rho || *eam_force ) && *eam_ghost ) *ghost_dist_max = std::max( *ghost_dist_max , (*rcut) * 2.0 );
      XsmdState& st = XsmdState::instance();
      if( ! st.on_gpu ) return;  // before the first neighbour list, as the reference's empty grid
      if( ! st.tables_loaded )
      {
        // The knots back from the spline rows eam_alloy_init built: row m + 1 of a
        // table holds the value at knot m in its coefficient 6 (eam_alloy.cpp:29-58),
        // and the library rebuilds the same coefficients from them.
        const auto& p = *parameters;
        auto knots = []( const std::vector< std::vector< std::vector<double> > >& tab, int n )
        {
          std::vector<double> k;
          for( const auto& t : tab ) for( int m = 0; m < n; m++ ) k.push_back( t[m + 1][6] );
          return k;
        };
        const std::vector<double> frho = knots( p.frho_spline, p.nrho ), rhor = knots( p.rhor_spline, p.nr ),
                                  z2r = knots( p.z2r_spline, p.nr );
        st.check( xs_eam_alloy_tables( st.ctx, p.nelements, p.nrho, 1.0 / p.rdrho, p.nr, 1.0 / p.rdr, p.rc,
                                       frho.data(), rhor.data(), z2r.data() ),
                  "xs_eam_alloy_tables" );
        st.tables_loaded = true;
      }
      // types slot: a species missing from a non-empty list switches off its pairs
      unsigned enabled = ~0u;
      if( ! types->empty() )
      {
        enabled = 0u;
        for( const auto& name : *types )
          for( size_t s = 0; s < species->size(); s++ )
            if( species->at(s).name() == name ) enabled |= 1u << s;
      }
      st.check( xs_eam_alloy_types( st.ctx, enabled ), "xs_eam_alloy_types" );
      unsigned flags = *mixed_precision ? XS_MIXED : 0u;
      if( trigger_thermo_state.has_value() && *trigger_thermo_state ) flags |= XS_ENERGY | ( *compute_virial ? XS_VIRIAL : 0u );
      if( *eam_symmetry ) flags |= XS_SYMMETRIC;
      st.check( xs_eam_alloy_force( st.ctx, *rcut, *eam_rho, *eam_rho2emb, *eam_force, flags ), "xs_eam_alloy_force" );
    }
  };

  // lj_compute_force (pair_potential_impl.hxx:39-513): slots of :108-121
  template<class GridT>
  class XsmdLjComputeForce : public OperatorNode
  {
    ADD_SLOT( LennardJonesParms , parameters , INPUT , LennardJonesParms{} );
    ADD_SLOT( double            , rcut       , INPUT , REQUIRED );
    ADD_SLOT( double            , rcut_max   , INPUT_OUTPUT , 0.0 );
    ADD_SLOT( ParticleSpecies   , species    , INPUT , OPTIONAL );
    ADD_SLOT( std::string       , type       , INPUT , OPTIONAL );
    ADD_SLOT( bool              , ghost      , INPUT , false );
    ADD_SLOT( GridT             , grid       , INPUT_OUTPUT );
    ADD_SLOT( bool , trigger_thermo_state , INPUT , OPTIONAL );
    ADD_SLOT( bool , compute_virial , INPUT , true );
    ADD_SLOT( bool , mixed_precision , INPUT , true );
  public:
    inline void execute() override final
    {
      *rcut_max = std::max( *rcut , *rcut_max );
      if( species.has_value() && species->size() > 1 && ! type.has_value() )
        fatal_error() << "Error: exactly 1 atom type expected (single material) but found " << species->size() << std::endl;
      XsmdState& st = XsmdState::instance();
      if( ! st.on_gpu ) return;
      unsigned flags = *mixed_precision ? XS_MIXED : 0u;
      if( ! trigger_thermo_state.has_value() || *trigger_thermo_state ) flags |= XS_ENERGY | ( *compute_virial ? XS_VIRIAL : 0u );
      st.check( xs_lj_compute_force( st.ctx, parameters->epsilon, parameters->sigma, *rcut, flags ), "xs_lj_compute_force" );
    }
  };

  // zero_force_energy (zero_force_energy.cu:38-147), force_to_accel (force_to_accel.cu:38-122)
  class XsmdZeroForceEnergy : public OperatorNode
  {
  public:
    inline void execute() override final
    {
      XsmdState& st = XsmdState::instance();
      if( st.on_gpu ) st.check( xs_zero_force_energy( st.ctx ), "xs_zero_force_energy" );
    }
  };
  class XsmdForceToAccel : public OperatorNode
  {
  public:
    inline void execute() override final
    {
      XsmdState& st = XsmdState::instance();
      if( st.on_gpu ) st.check( xs_force_to_accel( st.ctx ), "xs_force_to_accel" );
    }
  };

  // push_f_v_r / push_f_v (config_numerical_schemes.msp:23-28): dt * dt_scale;
  // push_f_v_r's xform_mode INV_XFORM maps the step by the inverse domain xform
  class XsmdPushFVR : public OperatorNode
  {
    ADD_SLOT( double      , dt         , INPUT , REQUIRED );
    ADD_SLOT( double      , dt_scale   , INPUT , 1.0 );
    ADD_SLOT( std::string , xform_mode , INPUT , std::string("INV_XFORM") );
  public:
    inline void execute() override final
    {
      XsmdState& st = XsmdState::instance();
      const int mode = ( *xform_mode == "IDENTITY" ) ? XS_XFORM_IDENTITY : XS_XFORM_INV;
      if( st.on_gpu ) st.check( xs_push_f_v_r( st.ctx, *dt, *dt_scale, mode ), "xs_push_f_v_r" );
    }
  };
  class XsmdPushFV : public OperatorNode
  {
    ADD_SLOT( double , dt       , INPUT , REQUIRED );
    ADD_SLOT( double , dt_scale , INPUT , 1.0 );
  public:
    inline void execute() override final
    {
      XsmdState& st = XsmdState::instance();
      if( st.on_gpu ) st.check( xs_push_f_v( st.ctx, *dt, *dt_scale ), "xs_push_f_v" );
    }
  };

  // simulation_thermodynamic_state (simulation_thermodynamic_state.cpp:57-239): the
  // 27-double reduction on the GPU, then the reference's post-reduction steps
  // (:196-228; one process, identity transform)
  class XsmdThermodynamicState : public OperatorNode
  {
    ADD_SLOT( Domain             , domain                 , INPUT , REQUIRED );
    ADD_SLOT( double             , potential_energy_shift , INPUT , 0.0 );
    ADD_SLOT( ThermodynamicState , thermodynamic_state    , OUTPUT );
  public:
    inline void execute() override final
    {
      XsmdState& st = XsmdState::instance();
      double s[27] = {};
      if( st.on_gpu ) st.check( xs_thermo( st.ctx, s ), "xs_thermo" );
      using onika::math::Mat3d;
      using onika::math::Vec3d;
      const Mat3d virial{ s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7], s[8] };
      const Mat3d ke_tensor{ s[9], s[10], s[11], s[12], s[13], s[14], s[15], s[16], s[17] };
      const Vec3d momentum{ s[18], s[19], s[20] };
      const Vec3d kinetic_energy{ s[21], s[22], s[23] };
      const double mass = s[25];
      const Vec3d temperature = 2. * ( kinetic_energy - 0.5 * momentum * momentum / mass );
      const Mat3d ke_test = 2. * ( ke_tensor - 0.5 * tensor( momentum, momentum ) / mass );
      const double volume = bounds_volume( domain->bounds() );
      const Vec3d phydro = ( Vec3d{ ke_test.m11, ke_test.m22, ke_test.m33 } + Vec3d{ virial.m11, virial.m22, virial.m33 } ) / volume;
      const Vec3d pdev = ( Vec3d{ ke_test.m12, ke_test.m13, ke_test.m23 } + Vec3d{ virial.m12, virial.m13, virial.m23 } ) / volume;
      ThermodynamicState& t = *thermodynamic_state;
      t.set_virial( virial );
      t.set_ke_tensor( ke_test );
      t.set_pressure( phydro );
      t.set_deviator( pdev );
      t.set_kinetic_energy( kinetic_energy );
      t.set_temperature( temperature );
      t.set_kinetic_momentum( momentum );
      t.set_potential_energy( s[24] + *potential_energy_shift );
      t.set_internal_energy( 0. );
      t.set_chemical_energy( 0. );
      t.set_mass( mass );
      t.set_volume( volume );
      t.set_particle_count( static_cast<size_t>( s[26] ) );
    }
  };

  // the GPU state back into the grid, for operators that read the grid itself
  template<class GridT>
  class XsmdSyncGrid : public OperatorNode
  {
    ADD_SLOT( GridT , grid , INPUT_OUTPUT );
  public:
    inline void execute() override final
    {
      XsmdState& st = XsmdState::instance();
      if( ! st.on_gpu ) return;
      xsmd_download( st, *grid, grid->has_allocated_field( field::ep ), grid->has_allocated_field( field::virial ) );
    }
  };

  ONIKA_AUTORUN_INIT(xsmd_operators)
  {
    OperatorNodeFactory::instance()->register_factory( "xsmd_chunk_neighbors", make_grid_variant_operator< XsmdChunkNeighbors > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_particle_displ_over", make_simple_operator< XsmdParticleDisplOver > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_eam_alloy_force", make_grid_variant_operator< XsmdEamAlloyForce > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_lj_compute_force", make_grid_variant_operator< XsmdLjComputeForce > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_zero_force_energy", make_simple_operator< XsmdZeroForceEnergy > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_force_to_accel", make_simple_operator< XsmdForceToAccel > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_push_f_v_r", make_simple_operator< XsmdPushFVR > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_push_f_v", make_simple_operator< XsmdPushFV > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_simulation_thermodynamic_state", make_simple_operator< XsmdThermodynamicState > );
    OperatorNodeFactory::instance()->register_factory( "xsmd_sync_grid", make_grid_variant_operator< XsmdSyncGrid > );
  }
}
