! module_pepc_gpu.f90 -- Fortran host shim of the MI355X tree solve.
!
! Keeps PEPC's frontend-facing Fortran API (module_pepc, module_pepc.f90:32-56)
! and its Coulomb XYZQ particle type (module_pepc_types.f90:50-58,
! coulomb/module_interaction_specific_types_XYZQ.f90:30-42) so a pepc-mini or
! pepc-essential style frontend compiles against it unchanged; the treecode
! underneath is the HIP library behind include/pepc_gpu.h, bound here with
! iso_c_binding.  The shim copies the array-of-structures particles into
! structure-of-arrays buffers, calls the C-ABI, and reproduces the
! reference's side effects: grow reorders the particle array by key and
! resets %work to 1 (module_domains.f90:204-234, :318), traverse adds the
! fields and potentials to %results and counts interactions in %work.
! A nonzero status aborts with the library's message, as DEBUG_ERROR does
! (pepc_debug.h:28-35).

module module_pepc_kinds
   use iso_c_binding, only: c_double, c_int64_t, c_int32_t
   implicit none
   ! module_pepc_kinds.f90: 8-byte reals and integers for keys and counts
   integer, parameter :: kind_physics = c_double
   integer, parameter :: kind_particle = c_int64_t
   integer, parameter :: kind_key = c_int64_t
   integer, parameter :: kind_node = c_int64_t
   integer, parameter :: kind_default = c_int32_t
   integer, parameter :: kind_pe = kind_default
   integer, parameter :: kind_dim = 4
end module module_pepc_kinds

module module_interaction_specific_types
   use module_pepc_kinds
   implicit none
   type t_particle_data
      real(kind_physics) :: q            !< electric charge
   end type t_particle_data

   type t_particle_results
      real(kind_physics), dimension(3) :: e
      real(kind_physics) :: pot
   end type t_particle_results

   type(t_particle_results), parameter :: EMPTY_PARTICLE_RESULTS = &
      t_particle_results([0._kind_physics, 0._kind_physics, 0._kind_physics], 0._kind_physics)

   !> Node moments in the reference's layout (read back from the device by
   !> pepc_gpu_read_tree; declared for the frontends' VTK callbacks).
   type t_tree_node_interaction_data
      real(kind_physics) :: coc(3)
      real(kind_physics) :: charge
      real(kind_physics) :: abs_charge
      real(kind_physics) :: dip(3)
      real(kind_physics) :: quad(3)
      real(kind_physics) :: xyquad
      real(kind_physics) :: yzquad
      real(kind_physics) :: zxquad
      real(kind_physics) :: bmax
   end type t_tree_node_interaction_data
end module module_interaction_specific_types

module module_pepc_types
   use module_pepc_kinds
   use module_interaction_specific_types
   implicit none
   type t_particle
      real(kind_physics) :: x(1:3)        !< coordinates
      real*8 :: work                      !< work load (1 + interactions after traverse)
      integer(kind_key) :: key            !< space-filling key on the finest level
      integer(kind_node) :: node_leaf     !< not tracked by the GPU tree (0)
      integer(kind_particle) :: label     !< frontend label
      type(t_particle_data) :: data
      type(t_particle_results) :: results
   end type t_particle

   !> module_box.f90:33-38
   type t_box
      real(kind_physics) :: boxmin(3) = 0.0_kind_physics
      real(kind_physics) :: boxmax(3) = 0.0_kind_physics
      real(kind_physics) :: boxsize(3) = 0.0_kind_physics
   end type t_box

   !> The host-side record of the reference's tree (module_tree.f90:80-106).
   !> The nodes of this build live on the device: after a grow the record
   !> holds the tree's counts and bounding box and is marked allocated until
   !> pepc_timber_tree; it has no node array (the tree writers read the nodes
   !> from the device).
   type t_tree
      logical :: allocated = .false.
      integer(kind_particle) :: npart = 0, npart_me = 0
      integer(kind_node) :: nleaf = 0, ntwig = 0, nleaf_me = 0, ntwig_me = 0
      integer(kind_node) :: nbranch = 0, nbranch_me = 0
      type(t_box) :: bounding_box
   end type t_tree
end module module_pepc_types

!> Key algebra the frontends and the tree writers use
!> (module_spacefilling.f90:47-57, 445-565): level of a key and the inverse
!> Hilbert map to the cell's integer and real coordinates.
module module_spacefilling
   use module_pepc_kinds
   use module_pepc_types, only: t_box
   implicit none
   private
   public :: level_from_key, key_to_coord, key_to_intcoord_hilbert

   integer, parameter :: maxlevel = 20
   ! state tables of the inverse map: child value -> cell bits, exchange and
   ! reverse codes (5: swap x and z, 6: swap y and z)
   integer, parameter :: cval_of(0:7) = [0, 1, 3, 2, 6, 7, 5, 4]
   integer, parameter :: exch_of(0:7) = [5, 6, 0, 5, 5, 0, 6, 5]
   integer, parameter :: rev_of(0:7) = [0, 0, 0, 5, 0, 0, 6, 5]

contains

   !> position of the placeholder bit / 3
   integer function level_from_key(key)
      integer(kind_key), intent(in) :: key
      level_from_key = (63 - leadz(key)) / 3
   end function level_from_key

   !> integer coordinates (finest-cell units) of the lower corner of the
   !> key's cell
   subroutine key_to_intcoord_hilbert(key, ic)
      integer(kind_key), intent(in) :: key
      integer(kind_key), intent(out) :: ic(3)
      integer :: lev, i, j, k, h, cv, rv, ex
      integer(kind_key) :: t
      lev = level_from_key(key)
      ic = 0
      do i = 0, lev - 1
         h = int(iand(ishft(key, -3 * i), 7_kind_key))
         cv = cval_of(h)
         rv = rev_of(h)
         ex = exch_of(h)
         if (i > 0) then
            do j = 0, 2
               if (btest(rv, j)) ic(j + 1) = iand(not(ic(j + 1)), ishft(1_kind_key, i) - 1)
            end do
            do j = 0, 1
               if (.not. btest(ex, j)) cycle
               do k = j + 1, 2
                  if (btest(ex, k)) then
                     t = ic(j + 1)
                     ic(j + 1) = ic(k + 1)
                     ic(k + 1) = t
                  end if
               end do
            end do
         end if
         do j = 0, 2
            if (btest(cv, j)) ic(j + 1) = ior(ic(j + 1), ishft(1_kind_key, i))
         end do
      end do
      ic = ishft(ic, maxlevel - lev)
   end subroutine key_to_intcoord_hilbert

   !> module_spacefilling.f90:445-471: ((ic + 1/2) s + boxmin, s = boxsize / 2^20
   subroutine key_to_coord(b, key, x)
      type(t_box), intent(in) :: b
      integer(kind_key), intent(in) :: key
      real(kind_physics), intent(inout) :: x(3)
      integer(kind_key) :: ic(3)
      real(kind_physics) :: s(3)
      call key_to_intcoord_hilbert(key, ic)
      s = b%boxsize / 2_kind_key**maxlevel
      x = (real(ic, kind_physics) + 0.5_kind_physics) * s + b%boxmin
   end subroutine key_to_coord
end module module_spacefilling

!> iso_c_binding interface of include/pepc_gpu.h
module pepc_gpu_c
   use iso_c_binding
   implicit none

   type, bind(C) :: pepc_gpu_config
      real(c_double) :: theta2, eps2, sign, node_factor
      integer(c_int32_t) :: mac_select, group_size, precision, force_cubic_domain
      integer(c_int32_t) :: walk_blocks, device, group_pack, list_tile
   end type pepc_gpu_config

   type, bind(C) :: pepc_gpu_periodic
      integer(c_int32_t) :: periodicity(3), mirror_box_layers
      real(c_double) :: t_lattice(9), origin(3)
      integer(c_int32_t) :: far_field, extrinsic_correction
   end type pepc_gpu_periodic

   type, bind(C) :: pepc_gpu_stats
      integer(c_int64_t) :: n, nnodes, nleaf, ntwig, ngroups, dup_fixed
      integer(c_int64_t) :: interactions, mac_evaluations, p2p, m2p, amb_nodes, walk_steps
      integer(c_int32_t) :: flags, maxlevel
      integer(c_int32_t) :: level_offset(22)
   end type pepc_gpu_stats

   interface
      subroutine pepc_gpu_config_default(cfg) bind(C, name="pepc_gpu_config_default")
         import :: pepc_gpu_config
         type(pepc_gpu_config), intent(out) :: cfg
      end subroutine

      integer(c_int) function pepc_gpu_create(ctx, cfg) bind(C, name="pepc_gpu_create")
         import :: c_int, c_ptr, pepc_gpu_config
         type(c_ptr), intent(out) :: ctx
         type(pepc_gpu_config), intent(in) :: cfg
      end function

      integer(c_int) function pepc_gpu_destroy(ctx) bind(C, name="pepc_gpu_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function pepc_gpu_last_error(ctx, buf, len) bind(C, name="pepc_gpu_last_error")
         import :: c_int, c_ptr, c_char, c_size_t
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_size_t), value :: len
      end function

      integer(c_int) function pepc_gpu_reserve_nodes(ctx, nodes) bind(C, name="pepc_gpu_reserve_nodes")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nodes
      end function

      integer(c_int) function pepc_gpu_grow(ctx, n, x, y, z, q, keys, perm, xs, ys, zs) &
         bind(C, name="pepc_gpu_grow")
         import :: c_int, c_ptr, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: x(*), y(*), z(*), q(*)
         integer(c_int64_t), intent(out) :: keys(*), perm(*)
         real(c_double), intent(out) :: xs(*), ys(*), zs(*)
      end function

      integer(c_int) function pepc_gpu_traverse(ctx, ex, ey, ez, pot, work) &
         bind(C, name="pepc_gpu_traverse")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: ex(*), ey(*), ez(*), pot(*), work(*)
      end function

      integer(c_int) function pepc_gpu_read_tree(ctx, cap, link, first, mom, nnodes) &
         bind(C, name="pepc_gpu_read_tree")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: cap
         type(c_ptr), value :: link, first, mom
         integer(c_int64_t), intent(out) :: nnodes
      end function

      integer(c_int) function pepc_gpu_read_sorted(ctx, keys, perm, x, y, z, q) &
         bind(C, name="pepc_gpu_read_sorted")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, keys, perm, x, y, z, q
      end function

      integer(c_int) function pepc_gpu_read_box(ctx, sys, boxmin, boxmax, boxsize) &
         bind(C, name="pepc_gpu_read_box")
         import :: c_int, c_ptr, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: sys
         real(c_double), intent(out) :: boxmin(3), boxmax(3), boxsize(3)
      end function

      integer(c_int) function pepc_gpu_sync(ctx, st) bind(C, name="pepc_gpu_sync")
         import :: c_int, c_ptr, pepc_gpu_stats
         type(c_ptr), value :: ctx
         type(pepc_gpu_stats), intent(out) :: st
      end function

      integer(c_int) function pepc_gpu_timber(ctx) bind(C, name="pepc_gpu_timber")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function pepc_gpu_timings(ctx, tg, tw) bind(C, name="pepc_gpu_timings")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: tg, tw
      end function

      subroutine pepc_gpu_periodic_default(p) bind(C, name="pepc_gpu_periodic_default")
         import :: pepc_gpu_periodic
         type(pepc_gpu_periodic), intent(out) :: p
      end subroutine

      integer(c_int) function pepc_gpu_set_periodic(ctx, p) bind(C, name="pepc_gpu_set_periodic")
         import :: c_int, c_ptr, pepc_gpu_periodic
         type(c_ptr), value :: ctx
         type(pepc_gpu_periodic), intent(in) :: p
      end function

      integer(c_int) function pepc_gpu_check_tree(ctx, nbad) bind(C, name="pepc_gpu_check_tree")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: nbad
      end function

      integer(c_int) function pepc_gpu_traverse_sinks(ctx, m, x, y, z, mode, ex, ey, ez, pot, work) &
         bind(C, name="pepc_gpu_traverse_sinks")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: m
         real(c_double), intent(in) :: x(*), y(*), z(*)
         integer(c_int32_t), value :: mode
         real(c_double), intent(out) :: ex(*), ey(*), ez(*), pot(*), work(*)
      end function

      integer(c_int) function pepc_gpu_direct(ctx, n, x, y, z, q, ntest, testidx, eps2, sign, &
                                              ex, ey, ez, pot) bind(C, name="pepc_gpu_direct")
         import :: c_int, c_ptr, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n, ntest
         real(c_double), intent(in) :: x(*), y(*), z(*), q(*)
         integer(c_int64_t), intent(in) :: testidx(*)
         real(c_double), value :: eps2, sign
         real(c_double), intent(out) :: ex(*), ey(*), ez(*), pot(*)
      end function
   end interface
end module pepc_gpu_c

!> Timers of the library and the frontends (lpepcsrc/module_timings.f90): the
!> same ids and semantics (a running timer holds minus its start time) and
!> timings_GatherAndOutput writing the timing/*.dat tables; this shim runs as
!> one process, so maximum, average and minimum coincide.  The library fills
!> t_fields_tree (grow), t_fields_passes (walk) and t_all (both) with the
!> device times of each solve.
module module_timings
   use iso_c_binding
   implicit none
   private
   integer, parameter, public :: t_domains = 1, t_allocate = 2, t_restore = 7, t_walk = 8, &
      t_deallocate = 11, t_all = 12, t_local = 13, t_exchange_branches = 14, t_global = 15, &
      t_lattice = 16, t_fields_tree = 18, t_fields_passes = 20, t_fields_stats = 21, &
      t_domains_keys = 23, t_domains_sort = 24, t_domains_sort_pure = 25, t_domains_ship = 26, &
      t_domains_bound = 27, t_grow_tree = 29, t_build_pure = 34, t_props_leaves = 37, &
      t_tot = 41, t_direct_force = 49, t_direct_comm = 50, &
      t_userdefined_first = 60, t_userdefined_last = 90
   integer, parameter :: numtimings = t_userdefined_last
   real*8 :: tim(numtimings) = 0.0d0
   public :: timer_start, timer_stop, timer_resume, timer_add, timer_read, timer_reset
   public :: timer_reset_all, timings_GatherAndOutput

   interface
      integer(c_int) function timing_mkdir(path, mode) bind(C, name='mkdir')
         import :: c_int, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: mode
      end function timing_mkdir
   end interface

contains

   real*8 function wtime()
      integer(8) :: count, rate
      call system_clock(count, rate)
      wtime = real(count, 8) / real(rate, 8)
   end function wtime

   real*8 function timer_read(id)
      integer, intent(in) :: id
      timer_read = tim(id)
   end function timer_read

   subroutine timer_reset(id)
      integer, intent(in) :: id
      tim(id) = 0.0d0
   end subroutine timer_reset

   subroutine timer_reset_all()
      tim = 0.0d0
   end subroutine timer_reset_all

   subroutine timer_start(id)
      integer, intent(in) :: id
      tim(id) = -wtime()
   end subroutine timer_start

   subroutine timer_stop(id)
      integer, intent(in) :: id
      tim(id) = tim(id) + wtime()
   end subroutine timer_stop

   subroutine timer_resume(id)
      integer, intent(in) :: id
      tim(id) = tim(id) - wtime()
   end subroutine timer_resume

   subroutine timer_add(id, val)
      integer, intent(in) :: id
      real*8, intent(in) :: val
      tim(id) = tim(id) + val
   end subroutine timer_add

   subroutine table(itime, flag, t, filename, header)
      integer, intent(in) :: itime, flag
      real*8, intent(in) :: t(numtimings)
      character(*), intent(in) :: filename
      logical, intent(in) :: header
      integer :: u, i
      character(40) :: fmt
      open (newunit=u, file=filename, status='unknown', position='append')
      if (header) then
         write (fmt, '(a,i0,a)') '(2(1x,a20),', numtimings, '(1x,i20))'
         write (u, fmt) "#timestep", "userflag", [(i, i=1, numtimings)]
      end if
      write (fmt, '(a,i0,a)') '(2(1x,i20),', numtimings, '(1x,e20.5))'
      write (u, fmt) itime, flag, t
      close (u)
   end subroutine table

   !> Appends this step's timers to timing/timing_{max,avg,min,dev_abs,dev_rel}.dat
   !> (a header when itime <= 1 unless printheader says otherwise) and prints
   !> the total.
   subroutine timings_GatherAndOutput(itime, iuserflag, printheader)
      integer, intent(in) :: itime
      integer, optional, intent(in) :: iuserflag
      logical, optional, intent(in) :: printheader
      integer :: flag
      logical :: header
      integer(c_int) :: rc
      real*8 :: zero(numtimings)
      flag = 0
      if (present(iuserflag)) flag = iuserflag
      header = itime <= 1
      if (present(printheader)) header = printheader
      rc = timing_mkdir("timing"//c_null_char, int(o'755', c_int))
      zero = 0.0d0
      call table(itime, flag, tim, 'timing/timing_max.dat', header)
      call table(itime, flag, tim, 'timing/timing_avg.dat', header)
      call table(itime, flag, tim, 'timing/timing_min.dat', header)
      call table(itime, flag, zero, 'timing/timing_dev_abs.dat', header)
      call table(itime, flag, zero, 'timing/timing_dev_rel.dat', header)
      write (*, '(a20,f16.10," s")') "t_all = ", tim(t_all)
   end subroutine timings_GatherAndOutput
end module module_timings

!> Debug switches of the library (lpepcsrc/module_debug.f90:34-58): the
!> bitmask debug_level (from pepc_initialize's db_level_in or &libpepc) and
!> dbg(flag) to test it.
module module_debug
   implicit none
   private
   integer, public :: debug_level = 0
   integer, parameter, public :: DBG_STATUS = 1, DBG_TREE = 2, DBG_BUILD = 4, DBG_DOMAIN = 8, &
      DBG_BRANCH = 16, DBG_STATS = 32, DBG_WALKSUMMARY = 64, DBG_DUMPTREE = 128, &
      DBG_TIMINGFILE = 256, DBG_WALK = 1024, DBG_PERIODIC = 2048
   public :: dbg
contains
   logical function dbg(flag)
      integer, intent(in) :: flag
      dbg = iand(debug_level, flag) /= 0
   end function dbg
end module module_debug

!> The Coulomb backend's parameters (coulomb/module_interaction_specific.f90:
!> 35-43, the &calc_force_coulomb namelist).
module module_interaction_specific
   use module_pepc_kinds
   implicit none
   private
   logical, public :: include_far_field_if_periodic = .true.
   integer, public :: force_law = 3
   integer, public :: mac_select = 0
   real(kind_physics), public :: theta2 = 0.36_kind_physics
   real(kind_physics), public :: eps2 = 0.0_kind_physics
   ! read for the Kelbg law (force_law 6), which this backend does not offer
   real(kind_physics), public :: kelbg_invsqrttemp = 0.0_kind_physics
end module module_interaction_specific

!> Periodic boundaries (module_mirror_boxes.f90:30-36, 174-210): the
!> settings, the interaction cutoff (never cuts: huge, the reference's
!> default) and constrain_periodic for a host particle array.
module module_mirror_boxes
   use module_pepc_kinds
   use module_pepc_types
   implicit none
   private
   logical, public :: periodicity(3) = [.false., .false., .false.]
   integer, public :: mirror_box_layers = 1
   real(kind_physics), public :: t_lattice_1(3) = [1, 0, 0]
   real(kind_physics), public :: t_lattice_2(3) = [0, 1, 0]
   real(kind_physics), public :: t_lattice_3(3) = [0, 0, 1]
   real(kind_physics), public :: LatticeOrigin(3) = [0, 0, 0]
   real(kind_physics), public :: spatial_interaction_cutoff(3) = huge(0._kind_physics)
   public :: constrain_periodic
contains
   !> Wraps positions into the central box (module_mirror_boxes.f90:178-215).
   !> A lattice with positive axes along x, y, z (module_mirror_boxes.f90:
   !> 232-244): per periodic axis, Fortran MODULO relative to LatticeOrigin.
   !> Any other lattice: a particle with a lattice coordinate outside [0, 1]
   !> has all three wrapped, and only its periodic Cartesian components take
   !> the wrapped position.
   subroutine constrain_periodic(particles)
      type(t_particle), intent(inout) :: particles(:)
      real(kind_physics) :: a(3, 3), ainv(3, 3), det, c(3), r(3)
      integer(kind_particle) :: i
      integer :: k
      logical :: simple
      if (.not. any(periodicity)) return
      a(:, 1) = t_lattice_1; a(:, 2) = t_lattice_2; a(:, 3) = t_lattice_3
      simple = t_lattice_1(1) > 0 .and. t_lattice_1(2) == 0 .and. t_lattice_1(3) == 0 .and. &
               t_lattice_2(1) == 0 .and. t_lattice_2(2) > 0 .and. t_lattice_2(3) == 0 .and. &
               t_lattice_3(1) == 0 .and. t_lattice_3(2) == 0 .and. t_lattice_3(3) > 0
      if (simple) then
         do i = 1, size(particles, kind=kind_particle)
            do k = 1, 3
               if (periodicity(k)) particles(i)%x(k) = LatticeOrigin(k) + &
                  modulo(particles(i)%x(k) - LatticeOrigin(k), a(k, k))
            end do
         end do
         return
      end if
      det = a(1, 1) * (a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3)) - a(1, 2) * (a(2, 1) * a(3, 3) - &
            a(3, 1) * a(2, 3)) + a(1, 3) * (a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2))
      ainv(1, 1) = (a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3)) / det
      ainv(1, 2) = (a(1, 3) * a(3, 2) - a(1, 2) * a(3, 3)) / det
      ainv(1, 3) = (a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2)) / det
      ainv(2, 1) = (a(2, 3) * a(3, 1) - a(2, 1) * a(3, 3)) / det
      ainv(2, 2) = (a(1, 1) * a(3, 3) - a(1, 3) * a(3, 1)) / det
      ainv(2, 3) = (a(2, 1) * a(1, 3) - a(1, 1) * a(2, 3)) / det
      ainv(3, 1) = (a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2)) / det
      ainv(3, 2) = (a(3, 1) * a(1, 2) - a(1, 1) * a(3, 2)) / det
      ainv(3, 3) = (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2)) / det
      do i = 1, size(particles, kind=kind_particle)
         c = matmul(ainv, particles(i)%x - LatticeOrigin)
         if (any(c > 1.0_kind_physics) .or. any(c < 0.0_kind_physics)) then
            r = LatticeOrigin + matmul(a, modulo(c, 1.0_kind_physics))
            where (periodicity) particles(i)%x = r
         end if
      end do
   end subroutine constrain_periodic
end module module_mirror_boxes

!> The far field's extrinsic-to-intrinsic correction, as the frontends set it
!> (coulomb/module_fmm_framework.f90:42-49; pepc-b's setup.f90:141 selects
!> Redlack-Grindlay).  MEASUREMENT is unimplemented in the reference and
!> refused here.
module module_fmm_framework
   implicit none
   private
   integer, public, parameter :: FMM_EXTRINSIC_CORRECTION_NONE = 0
   integer, public, parameter :: FMM_EXTRINSIC_CORRECTION_REDLACK = 1
   integer, public, parameter :: FMM_EXTRINSIC_CORRECTION_FICTCHARGE = 2
   integer, public, parameter :: FMM_EXTRINSIC_CORRECTION_MEASUREMENT = 3
   integer, public :: fmm_extrinsic_correction = FMM_EXTRINSIC_CORRECTION_FICTCHARGE
end module module_fmm_framework

!> The frontend API of module_pepc.f90:32-56 on top of the GPU library.
module module_pepc
   use iso_c_binding
   use module_pepc_kinds
   use module_pepc_types
   use pepc_gpu_c
   use module_timings
   use module_debug
   use module_interaction_specific
   use module_mirror_boxes
   use module_fmm_framework, only: fmm_extrinsic_correction
   implicit none
   private

   public :: pepc_initialize, pepc_finalize, pepc_prepare
   public :: pepc_read_parameters, pepc_read_parameters_from_file_name, pepc_write_parameters
   public :: pepc_read_parameters_from_first_argument, pepc_get_para_file
   public :: pepc_particleresults_clear, pepc_grow_tree, pepc_traverse_tree
   public :: pepc_timber_tree, pepc_restore_particles, pepc_grow_and_traverse
   public :: pepc_grow_and_traverse_for_others, pepc_check_sanity
   public :: pepc_statistics, pepc_last_stats, pepc_last_timings, pepc_direct_force
   public :: pepc_device_tree
   !> module_pepc.f90's public tree: counts and box of the grown tree
   !> (its nodes stay on the device).
   type(t_tree), public, target :: global_tree

   ! &calc_force_coulomb lives in module_interaction_specific, the periodic
   ! settings in module_mirror_boxes (re-exported here)
   public :: include_far_field_if_periodic, force_law, mac_select, theta2, eps2, kelbg_invsqrttemp
   public :: periodicity, mirror_box_layers, t_lattice_1, t_lattice_2, t_lattice_3, LatticeOrigin
   ! &libpepc entries that apply (module_libpepc_main.f90:44, module_box.f90:29)
   logical, public :: force_cubic_domain = .false.
   real, public :: np_mult = 2.5
   ! &walk_para_gpu: the GPU walk's own parameters
   integer, public :: group_size = 16         ! walk tile size (module_tree_walk_simple TILE_SIZE)
   integer, public :: group_pack = 0
   integer, public :: list_tile = 0           ! LDS M2P list entries (0 = default)
   character(len=8), public :: precision = 'fp32'
   real(kind_physics), public :: interaction_sign = 1.0_kind_physics
   logical, public :: print_stats = .false.   ! also: debug_level with DBG_STATS
   ! &libpepc entries read for compatibility: curve_type must be 1 (Hilbert,
   ! as the reference requires), idim 3; np_mult and the list length factor
   ! size the node pool as tree_create does (module_tree.f90:177-191); the
   ! load-balancing weights and the thread count have no role on one GPU
   integer :: curve_type = 1, weighted = 1, interaction_list_length_factor = 1, num_threads = 3
   integer :: idim = 3

   namelist /calc_force_coulomb/ force_law, mac_select, include_far_field_if_periodic, theta2, eps2, &
      kelbg_invsqrttemp
   namelist /libpepc/ debug_level, periodicity, np_mult, curve_type, force_cubic_domain, weighted, &
      interaction_list_length_factor, mirror_box_layers, num_threads, idim, print_stats
   namelist /walk_para_gpu/ group_size, group_pack, list_tile, precision, interaction_sign

   type(c_ptr) :: ctx = c_null_ptr
   type(c_ptr) :: grown_addr = c_null_ptr      ! first element of the grown array
   integer(c_int64_t), allocatable :: last_perm(:)
   integer :: my_rank_ = 0

contains

   subroutine check(rc, where)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: where
      character(kind=c_char) :: buf(512)
      character(len=512) :: msg
      integer :: i, ierr
      if (rc == 0) return
      ierr = pepc_gpu_last_error(ctx, buf, int(512, c_size_t))
      msg = ' '
      do i = 1, 512
         if (buf(i) == c_null_char) exit
         msg(i:i) = buf(i)
      end do
      write (*, '(a,a,a,i0,a,a)') 'PEPC-GPU error in ', where, ' (status ', rc, '): ', trim(msg)
      error stop 1
   end subroutine check

   !> module_pepc.f90:70-175 (no MPI on one GPU: rank 0 of 1)
   subroutine pepc_initialize(frontendname, my_rank, n_cpu, init_mpi, db_level_in, comm)
      character(*), intent(in) :: frontendname
      integer, intent(out) :: my_rank, n_cpu
      logical, intent(in) :: init_mpi
      integer, intent(in), optional :: db_level_in
      integer, intent(inout), optional :: comm
      my_rank = 0
      n_cpu = 1
      my_rank_ = 0
      write (*, '(a,a)') ' PEPC-GPU frontend: ', trim(frontendname)
      if (present(db_level_in)) debug_level = db_level_in   ! module_pepc.f90:162-164
      if (present(comm) .or. init_mpi) continue
   end subroutine pepc_initialize

   !> module_pepc.f90:230-244.  A group the file lacks keeps its defaults; a
   !> group with an unknown entry or a bad value stops the run, as the
   !> reference's plain namelist reads do.
   subroutine pepc_read_parameters(filehandle)
      integer, intent(in) :: filehandle
      integer :: ios
      character(len=256) :: msg
      rewind (filehandle)
      read (filehandle, NML=libpepc, iostat=ios, iomsg=msg)
      call bad_group('libpepc')
      rewind (filehandle)
      read (filehandle, NML=calc_force_coulomb, iostat=ios, iomsg=msg)
      call bad_group('calc_force_coulomb')
      rewind (filehandle)
      read (filehandle, NML=walk_para_gpu, iostat=ios, iomsg=msg)
      call bad_group('walk_para_gpu')
      if (curve_type /= 1) then       ! module_libpepc_main.f90:57-60
         write (*, *) 'PEPC-GPU: only curve_type = 1 (Hilbert) is supported'
         error stop 1
      end if
      if (idim /= 3) then
         write (*, *) 'PEPC-GPU: only idim = 3 is supported'
         error stop 1
      end if
   contains
      subroutine bad_group(name)
         character(*), intent(in) :: name
         if (ios <= 0) return
         write (*, '(a,a,a,a)') ' PEPC-GPU: error reading namelist &', name, ': ', trim(msg)
         error stop 1
      end subroutine bad_group
   end subroutine pepc_read_parameters

   !> module_pepc.f90:250-262: the namelists this library reads
   subroutine pepc_write_parameters(filehandle)
      integer, intent(in) :: filehandle
      write (filehandle, NML=calc_force_coulomb)
      write (filehandle, NML=walk_para_gpu)
      write (filehandle, NML=libpepc)
   end subroutine pepc_write_parameters

   subroutine pepc_read_parameters_from_file_name(filename)
      character(*), intent(in) :: filename
      integer, parameter :: u = 47
      open (u, file=filename, status='old', action='read')
      call pepc_read_parameters(u)
      close (u)
   end subroutine pepc_read_parameters_from_file_name

   subroutine pepc_read_parameters_from_first_argument(read_para_file, para_file)
      logical, optional, intent(out) :: read_para_file
      character(len=255), optional, intent(out) :: para_file
      character(len=255) :: f
      if (command_argument_count() >= 1) then
         call get_command_argument(1, f)
         call pepc_read_parameters_from_file_name(trim(f))
         if (present(read_para_file)) read_para_file = .true.
         if (present(para_file)) para_file = f
      else
         if (present(read_para_file)) read_para_file = .false.
      end if
   end subroutine pepc_read_parameters_from_first_argument

   !> module_pepc.f90:333-367: the first command-line argument names the
   !> parameter file (rank 0 reads it; one rank here, so nothing to broadcast)
   subroutine pepc_get_para_file(available, file_name, my_rank, comm)
      logical, intent(out) :: available
      character(len=255), intent(out) :: file_name
      integer(kind_pe), intent(in) :: my_rank
      integer, intent(in), optional :: comm
      available = .false.
      file_name = ' '
      if (my_rank == 0) then
         if (command_argument_count() /= 0) then
            call get_command_argument(1, file_name)
            available = .true.
            write (*, *) "found parameter file: ", file_name
         end if
      end if
      if (present(comm)) continue
   end subroutine pepc_get_para_file

   !> Node table size of tree_create (module_tree.f90:172-191) for n particles:
   !> max(30 nintmax + 4 n, 10^4) for np_mult > 0, else |np_mult| 10^4, with
   !> nintmax the list factor times the Hernquist estimate of
   !> get_number_of_interactions_per_particle
   !> (coulomb/module_interaction_specific.f90:231-243).  Stops, as the
   !> reference does, when the table could not hold one node per particle;
   !> returns the table's entry count (at least 2^15).
   integer(c_int64_t) function reference_node_table(n) result(entries)
      integer(c_int64_t), intent(in) :: n
      integer(c_int64_t) :: nintmax, maxaddress
      real(kind_physics) :: inv
      inv = max(theta2 / (35.0_8 * log(1.0_8 * real(max(n, 1_c_int64_t), 8))), 1.0_8 / real(max(n, 1_c_int64_t), 8))
      nintmax = int(1.0_8 / inv, c_int64_t) * interaction_list_length_factor
      if (np_mult > 0) then
         maxaddress = max(30_c_int64_t * nintmax + 4_c_int64_t * n, 10000_c_int64_t)
      else
         maxaddress = int(abs(real(np_mult, 8)) * 10000.0_8, c_int64_t)
      end if
      if (maxaddress <= n) then
         write (*, '(a,i0,a,i0,a)') ' PEPC-GPU: maxaddress = ', maxaddress, ' .le. npart = ', n, &
            '. You should increase np_mult.'
         error stop 1
      end if
      entries = max(maxaddress, 32768_c_int64_t)
   end function reference_node_table

   !> module_pepc.f90:269-288: creates the GPU context with the parameters read
   subroutine pepc_prepare(idim)
      integer(kind_dim), optional, intent(in) :: idim
      type(pepc_gpu_config) :: cfg
      type(pepc_gpu_periodic) :: per
      if (present(idim)) then
         if (idim /= 3) then
            write (*, *) 'PEPC-GPU: only idim = 3 is supported'
            error stop 1
         end if
      end if
      if (force_law /= 3) then
         write (*, *) 'PEPC-GPU: only force_law = 3 (3-D Coulomb)'
         error stop 1
      end if
      if (c_associated(ctx)) call check(pepc_gpu_destroy(ctx), 'pepc_prepare')
      call pepc_gpu_config_default(cfg)
      cfg%theta2 = theta2
      cfg%eps2 = eps2
      cfg%mac_select = mac_select
      cfg%sign = interaction_sign
      cfg%node_factor = max(1.0_c_double, real(abs(np_mult), c_double))
      cfg%group_size = group_size
      cfg%group_pack = group_pack
      cfg%list_tile = list_tile
      cfg%force_cubic_domain = merge(1, 0, force_cubic_domain)
      if (trim(precision) == 'fp64') then
         cfg%precision = 1
      else
         cfg%precision = 0
      end if
      call check(pepc_gpu_create(ctx, cfg), 'pepc_prepare')
      if (any(periodicity)) then   ! neighbour_boxes_prepare + fmm_framework_prepare
         call pepc_gpu_periodic_default(per)
         per%periodicity = merge(1, 0, periodicity)
         per%mirror_box_layers = mirror_box_layers
         per%t_lattice(1:3) = t_lattice_1
         per%t_lattice(4:6) = t_lattice_2
         per%t_lattice(7:9) = t_lattice_3
         per%origin = LatticeOrigin
         per%far_field = merge(1, 0, include_far_field_if_periodic)
         per%extrinsic_correction = fmm_extrinsic_correction
         call check(pepc_gpu_set_periodic(ctx, per), 'pepc_prepare (periodic)')
      end if
   end subroutine pepc_prepare

   subroutine pepc_finalize(comm)
      integer, intent(in), optional :: comm
      if (c_associated(ctx)) call check(pepc_gpu_destroy(ctx), 'pepc_finalize')
      ctx = c_null_ptr
      if (allocated(last_perm)) deallocate (last_perm)
      if (present(comm)) continue
   end subroutine pepc_finalize

   !> coulomb/module_interaction_specific.f90:282-288
   subroutine pepc_particleresults_clear(particles)
      type(t_particle), intent(inout) :: particles(:)
      particles(:)%results = EMPTY_PARTICLE_RESULTS
   end subroutine pepc_particleresults_clear

   !> module_pepc.f90:435-446: builds the tree and reorders particles by key
   subroutine pepc_grow_tree(particles)
      type(t_particle), allocatable, target, intent(inout) :: particles(:)
      real(c_double), allocatable :: x(:), y(:), z(:), q(:), xs(:), ys(:), zs(:)
      integer(c_int64_t), allocatable :: keys(:)
      type(t_particle), allocatable :: tmp(:)
      integer(c_int64_t) :: n, i
      if (.not. c_associated(ctx)) call pepc_prepare()
      n = size(particles, kind=c_int64_t)
      allocate (x(n), y(n), z(n), q(n), xs(n), ys(n), zs(n), keys(n))
      if (allocated(last_perm)) deallocate (last_perm)
      allocate (last_perm(n))
      do i = 1, n
         x(i) = particles(i)%x(1)
         y(i) = particles(i)%x(2)
         z(i) = particles(i)%x(3)
         q(i) = particles(i)%data%q
      end do
      call check(pepc_gpu_reserve_nodes(ctx, reference_node_table(n)), 'pepc_grow_tree')
      call check(pepc_gpu_grow(ctx, n, x, y, z, q, keys, last_perm, xs, ys, zs), 'pepc_grow_tree')
      call move_alloc(particles, tmp)
      allocate (particles(n))
      do i = 1, n
         particles(i) = tmp(last_perm(i) + 1)
         particles(i)%key = keys(i)
         particles(i)%x = [xs(i), ys(i), zs(i)]   ! duplicate-key fix-up may move a particle
         particles(i)%work = 1.0_8                ! module_domains.f90:318
         particles(i)%node_leaf = 0
      end do
      grown_addr = c_null_ptr
      if (n > 0) grown_addr = c_loc(particles(1))
      deallocate (tmp, x, y, z, q, xs, ys, zs, keys)
      call fill_global_tree(n)
   end subroutine pepc_grow_tree

   !> The host record of the grown tree: counts (tree_stats,
   !> module_tree.f90:622-699), the branch nodes of one rank (the root's
   !> children: find_local_branches, module_tree_grow.f90:283-363, with the
   !> whole key range as the rank's domain) and the bounding box.
   subroutine fill_global_tree(n)
      integer(c_int64_t), intent(in) :: n
      type(pepc_gpu_stats) :: st
      call check(pepc_gpu_sync(ctx, st), 'pepc_grow_tree')
      global_tree%npart = n
      global_tree%npart_me = n
      global_tree%nleaf = st%nleaf
      global_tree%nleaf_me = st%nleaf
      global_tree%ntwig = st%ntwig
      global_tree%ntwig_me = st%ntwig
      global_tree%nbranch = int(st%level_offset(3), kind_node) - int(st%level_offset(2), kind_node)
      global_tree%nbranch_me = global_tree%nbranch
      call check(pepc_gpu_read_box(ctx, 0_c_int64_t, global_tree%bounding_box%boxmin, &
                                   global_tree%bounding_box%boxmax, global_tree%bounding_box%boxsize), &
                 'pepc_grow_tree')
      global_tree%allocated = .true.
   end subroutine fill_global_tree

   !> The device tree of the last grow on the host, for the tree writers (the
   !> reference's writers walk its host nodes): per node i (BFS order, a
   !> twig's children contiguous) link(:, i) = {first child (twig) or
   !> particle (leaf), number of children (0 for a leaf), particles below,
   !> level}, all 0-based; first(i) its first particle; mom(:, i) the 16
   !> moments in the reference layout {coc(3), charge, abs_charge, dip(3),
   !> quad(3), xy, yz, zx, bmax, 0}; nkey(i) the node's key (its first
   !> particle's key cut to the node's level).
   subroutine pepc_device_tree(link, first, mom, nkey)
      integer(c_int32_t), allocatable, target, intent(out) :: link(:, :), first(:)
      real(c_double), allocatable, target, intent(out) :: mom(:, :)
      integer(kind_key), allocatable, intent(out) :: nkey(:)
      integer(c_int64_t), allocatable, target :: keys(:)
      integer(c_int64_t) :: nn, i, m
      if (.not. global_tree%allocated) then
         write (*, *) 'PEPC-GPU: pepc_device_tree without a grown tree'
         error stop 1
      end if
      nn = global_tree%nleaf + global_tree%ntwig
      m = global_tree%npart
      allocate (link(4, nn), first(nn), mom(16, nn), nkey(nn), keys(m))
      call check(pepc_gpu_read_tree(ctx, nn, c_loc(link), c_loc(first), c_loc(mom), i), &
                 'pepc_device_tree')
      call check(pepc_gpu_read_sorted(ctx, c_loc(keys), c_null_ptr, c_null_ptr, c_null_ptr, &
                                      c_null_ptr, c_null_ptr), 'pepc_device_tree')
      do i = 1, nn
         nkey(i) = ishft(keys(first(i) + 1), -3 * (20 - link(4, i)))
      end do
      deallocate (keys)
   end subroutine pepc_device_tree

   !> module_pepc.f90:459-468.  The grown particles (in their grown order) take
   !> the grown groups; any other array is walked particle by particle, as the
   !> reference walks particles without keys (module_tree_walk_simple.f90:155-159),
   !> and its interaction counts are added to %work.
   subroutine pepc_traverse_tree(particles)
      type(t_particle), target, intent(inout) :: particles(:)
      real(c_double), allocatable :: x(:), y(:), z(:), ex(:), ey(:), ez(:), pot(:), work(:)
      integer(c_int64_t) :: n, i
      logical :: grown
      n = size(particles, kind=c_int64_t)
      if (.not. allocated(last_perm)) then
         write (*, *) 'PEPC-GPU: pepc_traverse_tree before pepc_grow_tree'
         error stop 1
      end if
      if (n == 0) return
      grown = n == size(last_perm, kind=c_int64_t) .and. c_associated(grown_addr, c_loc(particles(1)))
      allocate (ex(n), ey(n), ez(n), pot(n), work(n))
      if (grown) then
         call check(pepc_gpu_traverse(ctx, ex, ey, ez, pot, work), 'pepc_traverse_tree')
      else
         allocate (x(n), y(n), z(n))
         do i = 1, n
            x(i) = particles(i)%x(1)
            y(i) = particles(i)%x(2)
            z(i) = particles(i)%x(3)
         end do
         call check(pepc_gpu_traverse_sinks(ctx, n, x, y, z, 0_c_int32_t, ex, ey, ez, pot, work), &
                    'pepc_traverse_tree (other particles)')
         deallocate (x, y, z)
      end if
      do i = 1, n
         particles(i)%results%e = particles(i)%results%e + [ex(i), ey(i), ez(i)]
         particles(i)%results%pot = particles(i)%results%pot + pot(i)
         ! the walk adds one per leaf and accepted twig to %work
         ! (module_tree_walk_simple.f90:335, :380); the grown particles' work
         ! from the library starts at the grow's 1
         if (grown) then
            particles(i)%work = particles(i)%work + (work(i) - 1.0_8)
         else
            particles(i)%work = particles(i)%work + work(i)
         end if
      end do
      deallocate (ex, ey, ez, pot, work)
      call record_timers()
   end subroutine pepc_traverse_tree

   !> The library's timers for this solve from the device clock: tree build
   !> (t_fields_tree), walk (t_fields_passes) and both (t_all), in seconds.
   subroutine record_timers()
      real(c_double) :: tg, tw
      call check(pepc_gpu_timings(ctx, tg, tw), 'pepc_traverse_tree')
      call timer_reset(t_fields_tree)
      call timer_reset(t_fields_passes)
      call timer_reset(t_all)
      if (tg > 0) call timer_add(t_fields_tree, real(tg, 8) * 1.0d-3)
      if (tw > 0) call timer_add(t_fields_passes, real(tw, 8) * 1.0d-3)
      call timer_add(t_all, timer_read(t_fields_tree) + timer_read(t_fields_passes))
   end subroutine record_timers

   !> module_pepc.f90:538
   subroutine pepc_timber_tree()
      if (c_associated(ctx)) call check(pepc_gpu_timber(ctx), 'pepc_timber_tree')
      global_tree%allocated = .false.
   end subroutine pepc_timber_tree

   !> module_pepc.f90:525 -> domain_restore, module_domains.f90:372-408
   subroutine pepc_restore_particles(particles)
      type(t_particle), allocatable, intent(inout) :: particles(:)
      type(t_particle), allocatable :: tmp(:)
      integer(c_int64_t) :: i
      if (.not. allocated(last_perm)) return
      call move_alloc(particles, tmp)
      allocate (particles(size(tmp)))
      do i = 1, size(tmp, kind=c_int64_t)
         particles(last_perm(i) + 1) = tmp(i)
      end do
      deallocate (tmp)
   end subroutine pepc_restore_particles

   !> module_pepc.f90:374-429
   subroutine pepc_grow_and_traverse(particles, itime, no_dealloc, no_restore)
      type(t_particle), allocatable, intent(inout) :: particles(:)
      integer, intent(in) :: itime
      logical, optional, intent(in) :: no_dealloc, no_restore
      logical :: restore, dealloc
      restore = .true.
      dealloc = .true.
      if (present(no_dealloc)) dealloc = .not. no_dealloc
      if (present(no_restore)) restore = .not. no_restore
      call pepc_grow_tree(particles)
      call pepc_traverse_tree(particles)
      if (print_stats .or. dbg(DBG_STATS)) call pepc_statistics(itime)
      if (restore) call pepc_restore_particles(particles)
      if (dealloc) call pepc_timber_tree()
   end subroutine pepc_grow_and_traverse

   !> module_pepc.f90:394-433: forces of particles_source on particles_sink
   subroutine pepc_grow_and_traverse_for_others(particles_source, particles_sink, itime, no_dealloc, &
                                                no_restore)
      type(t_particle), allocatable, target, intent(inout) :: particles_source(:)
      type(t_particle), allocatable, optional, target, intent(inout) :: particles_sink(:)
      integer, intent(in) :: itime
      logical, optional, intent(in) :: no_dealloc, no_restore
      logical :: restore, dealloc
      restore = .true.
      dealloc = .true.
      if (present(no_dealloc)) dealloc = .not. no_dealloc
      if (present(no_restore)) restore = .not. no_restore
      call pepc_grow_tree(particles_source)
      if (present(particles_sink)) then
         call pepc_traverse_tree(particles_sink)
      else
         call pepc_traverse_tree(particles_source)
      end if
      if (print_stats .or. dbg(DBG_STATS)) call pepc_statistics(itime)
      if (restore) call pepc_restore_particles(particles_source)
      if (dealloc) call pepc_timber_tree()
   end subroutine pepc_grow_and_traverse_for_others

   !> directforce (src/utils/module_directsum.f90:34-172) on the device, open
   !> boundaries: field and potential at particles(testidx(1:ntest)) from all
   !> particles, self excluded by index, with this module's eps2 and
   !> interaction sign.  The periodic form (mirror boxes and the lattice term)
   !> is not provided and stops with an error.
   subroutine pepc_direct_force(particles, testidx, ntest, directresults)
      type(t_particle), intent(in) :: particles(:)
      integer(kind_particle), intent(in) :: testidx(:)
      integer(kind_particle), intent(in) :: ntest
      type(t_particle_results), intent(out) :: directresults(:)
      real(c_double), allocatable :: x(:), y(:), z(:), q(:), ex(:), ey(:), ez(:), pot(:)
      integer(c_int64_t), allocatable :: idx(:)
      ! with periodicity, the mirror boxes and the lattice far field set in
      ! pepc_prepare apply (module_directsum.f90:80-93, 130-137)
      if (.not. c_associated(ctx)) then
         write (*, *) 'PEPC-GPU: directforce needs pepc_prepare first'
         error stop 1
      end if
      if (ntest < 1) return
      x = particles(:)%x(1)
      y = particles(:)%x(2)
      z = particles(:)%x(3)
      q = particles(:)%data%q
      idx = int(testidx(1:ntest) - 1, c_int64_t)
      allocate (ex(ntest), ey(ntest), ez(ntest), pot(ntest))
      call check(pepc_gpu_direct(ctx, size(particles, kind=c_int64_t), x, y, z, q, &
                                 int(ntest, c_int64_t), idx, real(eps2, c_double), &
                                 real(interaction_sign, c_double), ex, ey, ez, pot), 'directforce')
      directresults(1:ntest)%e(1) = ex
      directresults(1:ntest)%e(2) = ey
      directresults(1:ntest)%e(3) = ez
      directresults(1:ntest)%pot = pot
   end subroutine pepc_direct_force

   !> module_pepc.f90:504-520: aborts if the grown tree is inconsistent;
   !> dump prints the tree's counters (the reference dumps its hash table)
   subroutine pepc_check_sanity(caller, dump, particles)
      character(*), intent(in) :: caller
      logical, optional, intent(in) :: dump
      type(t_particle), optional, intent(in) :: particles(:)
      integer(c_int64_t) :: nbad
      type(pepc_gpu_stats) :: st
      logical :: do_dump
      do_dump = .false.
      if (present(dump)) do_dump = dump
      call check(pepc_gpu_check_tree(ctx, nbad), 'pepc_check_sanity')
      if (nbad /= 0 .or. do_dump) then
         st = pepc_last_stats()
         write (*, '(a,a,a,i0,a,3i14)') ' tree check (', trim(caller), '): ', nbad, &
            ' violations; nodes/leaves/twigs ', st%nnodes, st%nleaf, st%ntwig
         if (present(particles)) write (*, '(a,i0)') ' particles: ', size(particles)
      end if
      if (nbad /= 0) then
         write (*, *) 'PEPC-GPU: Sanity check failed, aborting!'
         error stop 1
      end if
   end subroutine pepc_check_sanity

   function pepc_last_stats() result(st)
      type(pepc_gpu_stats) :: st
      call check(pepc_gpu_sync(ctx, st), 'pepc_last_stats')
   end function pepc_last_stats

   subroutine pepc_last_timings(t_grow_ms, t_walk_ms)
      real(c_double), intent(out) :: t_grow_ms, t_walk_ms
      call check(pepc_gpu_timings(ctx, t_grow_ms, t_walk_ms), 'pepc_last_timings')
   end subroutine pepc_last_timings

   !> module_pepc.f90:478-500: tree and walk statistics (tree_stats,
   !> tree_walk_statistics) to stdout
   subroutine pepc_statistics(itime)
      integer, intent(in) :: itime
      type(pepc_gpu_stats) :: st
      st = pepc_last_stats()
      write (*, '(a,i8)') ' ######## STATISTICS step ', itime
      write (*, '(a,3i14)') ' nodes / leaves / twigs: ', st%nnodes, st%nleaf, st%ntwig
      write (*, '(a,i14)') ' walk groups:             ', st%ngroups
      write (*, '(a,2i16)') ' interactions / mac evals:', st%interactions, st%mac_evaluations
      write (*, '(a,2i16)') ' p2p / m2p:               ', st%p2p, st%m2p
   end subroutine pepc_statistics

end module module_pepc

!> Particle checkpoints in the reference's file formats
!> (src/utils/module_checkpoint.f90), without MPI: one process writes and
!> reads the whole particle array.  Every t_particle component of the Coulomb
!> XYZQ type is 8 bytes (module_pepc_kinds.f90:63-70), so a record is the 96
!> bytes {x(3), work, key, node_leaf, label, q, e(3), pot} that
!> MPI_TYPE_particle_vec describes (module_pepc_types.f90:107-147).
!>   MPI-IO file  <dir>/mpi/particle_<itime:i12.12>.mpi: a 128-byte header
!>                {int64 n_total, int32 itime, zeros}, then n_total records
!>   binary file  <dir>/binary/particle_<itime:i12.12>_<rank:i6.6>.bin: records
!>   ASCII file   <dir>/ascii/particle_<itime:i12.12>_<rank:i6.6>.dat: one
!>                list-directed line per particle
!> each with a parameter file <name>.h (pepc_write_parameters).
module module_checkpoint
   use iso_c_binding
   use module_pepc_kinds
   use module_pepc_types
   implicit none
   private

   integer, parameter :: filehandle = 91
   character(12), parameter :: directory = './particles'
   integer(c_int64_t), parameter :: max_header_size = 128

   public write_particles_ascii, write_particles_binary, write_particles_mpiio
   public read_particles_mpiio, read_particles_mpiio_from_filename

   interface
      integer(c_int) function c_mkdir(path, mode) bind(C, name="mkdir")
         import :: c_int, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: mode
      end function
   end interface

contains

   subroutine create_directory(dir)      ! module_utils.f90 create_directory
      character(*), intent(in) :: dir
      integer(c_int) :: rc
      rc = c_mkdir(trim(dir)//c_null_char, int(o'755', c_int))   ! an existing directory is fine
   end subroutine create_directory

   subroutine write_params(filename)
      use module_pepc, only: pepc_write_parameters
      character(*), intent(in) :: filename
      open (filehandle, file=trim(filename)//'.h', status='replace', action='write')
      call pepc_write_parameters(filehandle)
      close (filehandle)
   end subroutine write_params

   !> module_checkpoint.f90:43-82
   subroutine write_particles_ascii(my_rank, itime, dp, filename_out)
      integer(kind_pe), intent(in) :: my_rank
      integer(kind_default), intent(in) :: itime
      type(t_particle), intent(in), dimension(:) :: dp
      character(*), optional, intent(out) :: filename_out
      character(100) :: filename
      integer(kind_particle) :: i
      call create_directory(trim(directory))
      call create_directory(trim(directory)//'/ascii')
      write (filename, '(a,"/ascii/particle_",i12.12,"_",i6.6,".dat")') trim(directory), itime, my_rank
      open (filehandle, file=trim(filename), status='replace', action='write')
      do i = 1, size(dp, kind=kind(i))
         write (filehandle, *) dp(i)
      end do
      close (filehandle)
      if (my_rank == 0) call write_params(filename)
      if (present(filename_out)) filename_out = trim(filename)//'.h'
   end subroutine write_particles_ascii

   !> module_checkpoint.f90:84-121
   subroutine write_particles_binary(my_rank, itime, dp, filename)
      integer(kind_pe), intent(in) :: my_rank
      integer(kind_default), intent(in) :: itime
      type(t_particle), intent(in), dimension(:) :: dp
      character(*), intent(out) :: filename
      call create_directory(trim(directory))
      call create_directory(trim(directory)//'/binary')
      write (filename, '(a,"/binary/particle_",i12.12,"_",i6.6,".bin")') trim(directory), itime, my_rank
      open (filehandle, file=trim(filename), status='replace', access='stream', form='unformatted')
      write (filehandle) dp(1:size(dp))
      close (filehandle)
      if (my_rank == 0) call write_params(filename)
      filename = trim(filename)//'.h'
   end subroutine write_particles_binary

   !> module_checkpoint.f90:122-193 on one process
   subroutine write_particles_mpiio(comm, itime, n_total, dp, filename)
      integer, intent(in) :: comm
      integer(kind_default), intent(in) :: itime
      integer*8, intent(in) :: n_total
      type(t_particle), intent(in) :: dp(:)
      character(*), intent(out) :: filename
      integer(c_int8_t) :: pad(max_header_size - 12)
      if (size(dp, kind=8) /= n_total) then
         write (*, '(a,i0,a,i0)') 'PEPC-GPU: Invalid total particle number: size(dp) = ', size(dp, kind=8), &
            ' but n_total = ', n_total
         error stop 1
      end if
      call create_directory(trim(directory))
      call create_directory(trim(directory)//'/mpi')
      write (filename, '(a,"/mpi/particle_",i12.12,".mpi")') trim(directory), itime
      pad = 0_c_int8_t
      open (filehandle, file=trim(filename), status='replace', access='stream', form='unformatted')
      write (filehandle) n_total, int(itime, c_int32_t), pad
      write (filehandle, pos=max_header_size + 1) dp(1:size(dp))
      close (filehandle)
      call write_params(filename)
      filename = trim(filename)//'.h'
      if (comm == comm) continue
   end subroutine write_particles_mpiio

   !> module_checkpoint.f90:195-214
   subroutine read_particles_mpiio(itime_in, comm, itime, n_total, dp, filename, nparticles_local, &
                                   file_exists, noparams)
      integer(kind_default) :: itime_in
      integer, intent(in) :: comm
      integer(kind_default), intent(out) :: itime
      integer(kind_particle), intent(out) :: n_total
      type(t_particle), allocatable :: dp(:)
      character(*), intent(out) :: filename
      integer(kind_default), optional, intent(in) :: nparticles_local
      logical, optional, intent(out) :: file_exists
      logical, optional, intent(in) :: noparams
      write (filename, '(a,"/mpi/particle_",i12.12,".mpi")') trim(directory), itime_in
      call read_particles_mpiio_from_filename(comm, itime, n_total, dp, filename, nparticles_local, &
                                              file_exists, noparams)
      filename = trim(filename)//'.h'
   end subroutine read_particles_mpiio

   !> module_checkpoint.f90:216-300 on one process (all records to this one)
   subroutine read_particles_mpiio_from_filename(comm, itime, n_total, dp, filename, nparticles_local, &
                                                 file_exists, noparams)
      use module_pepc, only: pepc_read_parameters
      integer, intent(in) :: comm
      integer(kind_default), intent(out) :: itime
      integer(kind_particle), intent(out) :: n_total
      type(t_particle), allocatable :: dp(:)
      character(*), intent(in) :: filename
      integer(kind_default), optional, intent(in) :: nparticles_local
      logical, optional, intent(out) :: file_exists
      logical, optional, intent(in) :: noparams
      logical :: there, noparams_
      integer(c_int32_t) :: it32
      integer(kind_particle) :: np_local
      noparams_ = .false.
      if (present(noparams)) noparams_ = noparams
      inquire (file=trim(filename), exist=there)
      if (present(file_exists)) file_exists = there
      if (.not. there) then
         if (present(file_exists)) return
         write (*, '(3a)') 'PEPC-GPU: read_particles_mpiio_from_filename(filename=', trim(filename), &
            ') - file does not exist'
         error stop 1
      end if
      open (filehandle, file=trim(filename), status='old', access='stream', form='unformatted', action='read')
      read (filehandle) n_total, it32
      itime = it32
      np_local = n_total
      if (present(nparticles_local)) then
         if (int(nparticles_local, kind_particle) /= n_total) then
            write (*, '(a,i0,a,i0)') 'PEPC-GPU: Invalid total particle number: nparticles_local = ', &
               nparticles_local, ' but the data file says n_total = ', n_total
            error stop 1
         end if
      end if
      if (allocated(dp)) deallocate (dp)
      allocate (dp(1:np_local))
      if (np_local > 0) read (filehandle, pos=max_header_size + 1) dp
      close (filehandle)
      if (.not. noparams_) then
         open (filehandle, file=trim(filename)//'.h', action='read', status='old')
         call pepc_read_parameters(filehandle)
         close (filehandle)
      end if
      if (comm == comm) continue
   end subroutine read_particles_mpiio_from_filename
end module module_checkpoint

!> The frontends' direct-sum check (src/utils/module_directsum.f90:34-172).
module module_directsum
   use module_pepc_kinds
   use module_pepc_types
   use module_interaction_specific_types
   use module_pepc, only: pepc_direct_force
   implicit none
   private
   public :: directforce

contains

   !> Fields and potentials at particles(testidx(1:ntest)) from all particles
   !> (open boundaries; one process, so comm is accepted and unused).
   subroutine directforce(particles, testidx, ntest, directresults, comm)
      type(t_particle), intent(in) :: particles(:)
      integer(kind_particle), dimension(:), intent(in) :: testidx
      integer(kind_particle), intent(in) :: ntest
      type(t_particle_results), dimension(:), intent(out) :: directresults
      integer, intent(in) :: comm
      call pepc_direct_force(particles, testidx, ntest, directresults)
   end subroutine directforce
end module module_directsum
