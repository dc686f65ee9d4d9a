! pepc_mini_gpu.f90 -- a pepc-mini style frontend on the GPU treecode.
!
! Follows the timestep loop of src/frontends/pepc-mini/pepc.f90:55-100:
! clear results, grow, traverse, push (kick then drift,
! pepc-mini/module_helper.f90:205-222), timber.  It is written only against
! module_pepc, so the same source builds against the reference library.
!
! Usage:
!   pepc_mini_gpu params                      ! timestep run, &pepcmini in params
!   pepc_mini_gpu params in.bin out.bin       ! one grow+traverse of the particles
!                                             ! in in.bin, results to out.bin
!   pepc_mini_gpu params in.bin out.bin sinks.bin sinks_out.bin
!                                             ! then also traverse the particles of
!                                             ! sinks.bin (another array) through
!                                             ! the same tree
!   pepc_mini_gpu --checkpoint in.bin out.bin ! write the particles of in.bin as
!                                             ! MPI-IO and binary checkpoints of
!                                             ! step 7 under ./particles, read the
!                                             ! MPI-IO one back into out.bin (no GPU)
!   pepc_mini_gpu --direct params in.bin out.bin
!                                             ! directforce for every particle of
!                                             ! in.bin (input order): out.bin holds
!                                             ! int64 n, float64 ex,ey,ez,pot(n)
!   pepc_mini_gpu --vtk in.bin name           ! write the particles of in.bin as a
!                                             ! three-step VTK series (steps 0, 1, 2;
!                                             ! times 0, 0.5, 1) with work = 1 + step,
!                                             ! pot = q (step + 1), field = (x (step + 1),
!                                             ! -y, z + step); step 1 adds "L2 error" =
!                                             ! 2 pot, step 2 scales coordinates by 2;
!                                             ! plus the space curve and the tree
!                                             ! writers (which skip: no host tree)
!                                             ! (no GPU)
!   pepc_mini_gpu --twice params in.bin out.bin
!                                             ! the parameter file through
!                                             ! pepc_get_para_file (argument 1 is
!                                             ! the flag, so the file is read by
!                                             ! name), one grow, two traversals of
!                                             ! the grown array: out.bin holds
!                                             ! int64 n, label(n), float64
!                                             ! x,y,z,ex,ey,ez,pot(n) after both,
!                                             ! work(n) after the first and after
!                                             ! the second traversal
!   pepc_mini_gpu --solve-output params in.bin out.bin
!                                             ! grow + traverse, then the outputs
!                                             ! of a solved step 3: MPI-IO
!                                             ! checkpoint (read back into out.bin:
!                                             ! int64 n, itime, then the 96-byte
!                                             ! records), branches and leaves as
!                                             ! VTK boxes; after pepc_timber_tree
!                                             ! the leaves writer skips (step 4)
!   pepc_mini_gpu --library params in.bin out.bin
!                                             ! the library modules without a GPU:
!                                             ! read params, wrap the particles of
!                                             ! in.bin with constrain_periodic
!                                             ! (out.bin: int64 n, float64 x,y,z(n)),
!                                             ! time two steps into timing/*.dat
!                                             ! (user timer t_userdefined_first + 1
!                                             ! gets 1.25 s, then 0.5 s more) and
!                                             ! print the settings read; the
!                                             ! lattice comes from &lattice_test
!                                             ! (t1, t2, t3, origin) if present
! in.bin : int64 n, then float64 x(n), y(n), z(n), q(n)
! out.bin: int64 n, int64 label(n) (0-based input index), float64 x,y,z(n),
!          ex,ey,ez,pot,work(n), all in the grown (key) order.
! sinks.bin: int64 m, float64 x(m), y(m), z(m); sinks_out.bin: int64 m, float64
!          ex,ey,ez,pot,work(m) in the sinks' order (work starts at 1).
!> A frontend's VTK data-array callback (module procedure: an internal one
!> passed as an argument would need an executable stack for its trampoline).
module pepc_mini_vtk
   use module_interaction_specific_types
   use module_vtk
   use module_vtk_helpers
   implicit none
contains
   !> The default arrays plus one more.
   subroutine with_l2(d, r, vtkf)
      type(t_particle_data), intent(in) :: d(:)
      type(t_particle_results), intent(in) :: r(:)
      type(vtkfile_unstructured_grid), intent(inout) :: vtkf
      call vtk_write_particle_data_results(d, r, vtkf)
      call vtkf%write_data_array("L2 error", 2.0_8 * r(:)%pot)
   end subroutine with_l2
end module pepc_mini_vtk

program pepc_mini_gpu
   use module_timings
   use module_debug
   use module_mirror_boxes
   use module_pepc_kinds
   use module_pepc_types
   use module_pepc
   use module_checkpoint
   use module_directsum
   use module_vtk
   use module_vtk_helpers
   use pepc_mini_vtk
   implicit none

   type(t_particle), allocatable :: particles(:), sinks(:)
   integer :: my_rank, n_ranks, step, nt, ip
   integer(kind_particle) :: tnp, i, n
   real(kind_physics) :: dt, epot, ekin
   real(kind_physics), allocatable :: v(:, :), a(:)
   character(len=255) :: fin, fout, fsin, fsout
   integer(8) :: t0, t1, rate
   real(8) :: tg, tw
   namelist /pepcmini/ tnp, nt, dt

   if (command_argument_count() >= 3) then
      call get_command_argument(1, fin)
      if (trim(fin) == '--checkpoint') then
         call checkpoint_roundtrip()
         stop
      end if
      if (trim(fin) == '--vtk') then
         call vtk_series()
         stop
      end if
      if (trim(fin) == '--direct') then
         call direct_all()
         stop
      end if
      if (trim(fin) == '--solve-output') then
         call solve_output()
         stop
      end if
      if (trim(fin) == '--twice') then
         call traverse_twice()
         stop
      end if
      if (trim(fin) == '--library') then
         call library_modules()
         stop
      end if
   end if

   call pepc_initialize("pepc-mini-gpu", my_rank, n_ranks, .true.)
   call pepc_read_parameters_from_first_argument()
   call pepc_prepare(3_kind_dim)

   if (command_argument_count() >= 3) then
      call get_command_argument(2, fin)
      call get_command_argument(3, fout)
      open (11, file=trim(fin), access='stream', form='unformatted', status='old')
      read (11) n
      allocate (particles(n), a(n))
      read (11) a; particles(:)%x(1) = a
      read (11) a; particles(:)%x(2) = a
      read (11) a; particles(:)%x(3) = a
      read (11) a; particles(:)%data%q = a
      close (11)
      do i = 1, n
         particles(i)%label = i - 1
         particles(i)%work = 1.0_8
      end do
      call pepc_particleresults_clear(particles)
      call pepc_grow_tree(particles)
      call pepc_check_sanity('pepc_mini_gpu', .false.)
      call pepc_traverse_tree(particles)
      open (12, file=trim(fout), access='stream', form='unformatted', status='replace')
      write (12) n
      write (12) particles(:)%label
      write (12) particles(:)%x(1)
      write (12) particles(:)%x(2)
      write (12) particles(:)%x(3)
      write (12) particles(:)%results%e(1)
      write (12) particles(:)%results%e(2)
      write (12) particles(:)%results%e(3)
      write (12) particles(:)%results%pot
      write (12) particles(:)%work
      close (12)
      if (command_argument_count() >= 5) then
         call get_command_argument(4, fsin)
         call get_command_argument(5, fsout)
         open (11, file=trim(fsin), access='stream', form='unformatted', status='old')
         read (11) n
         deallocate (a)
         allocate (sinks(n), a(n))
         read (11) a; sinks(:)%x(1) = a
         read (11) a; sinks(:)%x(2) = a
         read (11) a; sinks(:)%x(3) = a
         close (11)
         sinks(:)%data%q = 0.0_8
         sinks(:)%work = 1.0_8
         call pepc_particleresults_clear(sinks)
         call pepc_traverse_tree(sinks)          ! another array: walked sink by sink
         open (12, file=trim(fsout), access='stream', form='unformatted', status='replace')
         write (12) n
         write (12) sinks(:)%results%e(1)
         write (12) sinks(:)%results%e(2)
         write (12) sinks(:)%results%e(3)
         write (12) sinks(:)%results%pot
         write (12) sinks(:)%work
         close (12)
      end if
      call pepc_statistics(0)
      call pepc_timber_tree()
      call pepc_finalize()
      stop
   end if

   ! timestep run: uniform cube, alternating charges, unit masses, at rest
   tnp = 32768
   nt = 10
   dt = 0.01_kind_physics
   if (command_argument_count() >= 1) then
      call get_command_argument(1, fin)
      open (13, file=trim(fin), status='old', action='read')
      read (13, NML=pepcmini, iostat=ip)
      close (13)
   end if
   allocate (particles(tnp), v(3, tnp))
   call init_particles(particles)
   v = 0.0_kind_physics
   call system_clock(count_rate=rate)
   do step = 0, nt
      call pepc_particleresults_clear(particles)
      call system_clock(t0)
      call pepc_grow_tree(particles)
      call pepc_traverse_tree(particles)
      call system_clock(t1)
      call pepc_last_timings(tg, tw)
      epot = 0.5_kind_physics*sum(particles(:)%data%q*particles(:)%results%pot)
      ! velocities follow the particles through the reordering by label
      call push(particles, v, dt)
      ekin = 0.5_kind_physics*sum(v*v)
      write (*, '(a,i5,a,es12.4,a,es12.4,a,f9.3,a,f9.3,a,f9.3)') ' step', step, '  Epot', epot, &
         '  Ekin', ekin, '  grow+walk [ms]', 1e3*real(t1 - t0, 8)/real(rate, 8), '  dev grow', tg, &
         '  dev walk', tw
      if (dbg(DBG_STATS)) call pepc_statistics(step)
      call pepc_timber_tree()
      call timings_GatherAndOutput(step, 0)      ! pepc-mini/pepc.f90:99
   end do
   call pepc_finalize()

contains

   subroutine checkpoint_roundtrip()
      type(t_particle), allocatable :: back(:)
      character(len=255) :: f1, f2, f3
      integer(kind_default) :: itime
      integer(kind_particle) :: ntot
      call get_command_argument(2, fin)
      call get_command_argument(3, fout)
      open (11, file=trim(fin), access='stream', form='unformatted', status='old')
      read (11) n
      allocate (particles(n), a(n))
      read (11) a; particles(:)%x(1) = a
      read (11) a; particles(:)%x(2) = a
      read (11) a; particles(:)%x(3) = a
      read (11) a; particles(:)%data%q = a
      close (11)
      do i = 1, n
         particles(i)%work = 0.5_8*real(i, 8)
         particles(i)%key = 1000_8 + i
         particles(i)%node_leaf = i
         particles(i)%label = i - 1
         particles(i)%results%e = [real(i, 8), 2.0_8*real(i, 8), 3.0_8*real(i, 8)]
         particles(i)%results%pot = -real(i, 8)
      end do
      call write_particles_mpiio(0, 7, int(n, 8), particles, f1)
      call write_particles_binary(0, 7, particles, f2)
      call read_particles_mpiio(7, 0, itime, ntot, back, f3, noparams=.true.)
      open (12, file=trim(fout), access='stream', form='unformatted', status='replace')
      write (12) ntot, int(itime, 8)
      write (12) back
      close (12)
   end subroutine checkpoint_roundtrip

   subroutine solve_output()
      type(t_particle), allocatable :: back(:)
      character(len=255) :: f1, f3
      integer(kind_default) :: itime
      integer(kind_particle) :: ntot
      call pepc_initialize("pepc-mini-gpu", my_rank, n_ranks, .true.)
      call get_command_argument(2, fin)
      call pepc_read_parameters_from_file_name(trim(fin))
      call pepc_prepare(3_kind_dim)
      call get_command_argument(3, fin)
      call get_command_argument(4, fout)
      open (11, file=trim(fin), access='stream', form='unformatted', status='old')
      read (11) n
      allocate (particles(n), a(n))
      read (11) a; particles(:)%x(1) = a
      read (11) a; particles(:)%x(2) = a
      read (11) a; particles(:)%x(3) = a
      read (11) a; particles(:)%data%q = a
      close (11)
      do i = 1, n
         particles(i)%label = i - 1
         particles(i)%work = 1.0_8
      end do
      call pepc_particleresults_clear(particles)
      call pepc_grow_tree(particles)
      call pepc_traverse_tree(particles)
      call write_particles_mpiio(0, 3, int(n, 8), particles, f1)
      call read_particles_mpiio(3, 0, itime, ntot, back, f3, noparams=.true.)
      call vtk_write_branches(3, 0.25_8, VTK_STEP_FIRST, global_tree)
      call vtk_write_leaves(3, 0.25_8, VTK_STEP_FIRST, global_tree)
      call pepc_timber_tree()
      call vtk_write_leaves(4, 0.5_8, VTK_STEP_LAST, global_tree)
      open (12, file=trim(fout), access='stream', form='unformatted', status='replace')
      write (12) ntot, int(itime, 8)
      write (12) back
      close (12)
      call pepc_finalize()
   end subroutine solve_output

   subroutine traverse_twice()
      logical :: avail
      character(len=255) :: pf
      real(8), allocatable :: w1(:)
      call pepc_initialize("pepc-mini-gpu", my_rank, n_ranks, .true.)
      call pepc_get_para_file(avail, pf, my_rank)
      if (.not. avail .or. trim(pf) /= '--twice') stop 2
      call get_command_argument(2, pf)
      call pepc_read_parameters_from_file_name(trim(pf))
      call pepc_prepare(3_kind_dim)
      call get_command_argument(3, fin)
      call get_command_argument(4, fout)
      open (11, file=trim(fin), access='stream', form='unformatted', status='old')
      read (11) n
      allocate (particles(n), a(n), w1(n))
      read (11) a; particles(:)%x(1) = a
      read (11) a; particles(:)%x(2) = a
      read (11) a; particles(:)%x(3) = a
      read (11) a; particles(:)%data%q = a
      close (11)
      do i = 1, n
         particles(i)%label = i - 1
         particles(i)%work = 7.0_8          ! grow resets it to 1
      end do
      call pepc_particleresults_clear(particles)
      call pepc_grow_tree(particles)
      call pepc_traverse_tree(particles)
      w1 = particles(:)%work
      call pepc_traverse_tree(particles)
      open (12, file=trim(fout), access='stream', form='unformatted', status='replace')
      write (12) n
      write (12) particles(:)%label
      write (12) particles(:)%x(1)
      write (12) particles(:)%x(2)
      write (12) particles(:)%x(3)
      write (12) particles(:)%results%e(1)
      write (12) particles(:)%results%e(2)
      write (12) particles(:)%results%e(3)
      write (12) particles(:)%results%pot
      write (12) w1
      write (12) particles(:)%work
      close (12)
      call pepc_timber_tree()
      call pepc_finalize()
   end subroutine traverse_twice

   subroutine direct_all()
      type(t_particle_results), allocatable :: res(:)
      integer(kind_particle), allocatable :: idx(:)
      call get_command_argument(2, fin)
      call pepc_initialize("pepc-mini-gpu", my_rank, n_ranks, .true.)
      call pepc_read_parameters_from_file_name(trim(fin))
      call pepc_prepare(3_kind_dim)
      call get_command_argument(3, fin)
      call get_command_argument(4, fout)
      open (11, file=trim(fin), access='stream', form='unformatted', status='old')
      read (11) n
      allocate (particles(n), a(n), res(n), idx(n))
      read (11) a; particles(:)%x(1) = a
      read (11) a; particles(:)%x(2) = a
      read (11) a; particles(:)%x(3) = a
      read (11) a; particles(:)%data%q = a
      close (11)
      idx = [(i, i=1, n)]
      call directforce(particles, idx, n, res, 0)
      open (12, file=trim(fout), access='stream', form='unformatted', status='replace')
      write (12) n
      write (12) res(:)%e(1)
      write (12) res(:)%e(2)
      write (12) res(:)%e(3)
      write (12) res(:)%pot
      close (12)
      call pepc_finalize()
   end subroutine direct_all

   subroutine library_modules()
      integer :: k, ios
      real(kind_physics) :: t1(3), t2(3), t3(3), origin(3)
      namelist /lattice_test/ t1, t2, t3, origin
      call pepc_initialize("pepc-mini-gpu", my_rank, n_ranks, .true., db_level_in=0)
      call get_command_argument(2, fin)
      call pepc_read_parameters_from_file_name(trim(fin))
      ! the frontend's lattice (frontends set module_mirror_boxes' variables)
      t1 = t_lattice_1; t2 = t_lattice_2; t3 = t_lattice_3; origin = LatticeOrigin
      open (13, file=trim(fin), status='old', action='read')
      read (13, NML=lattice_test, iostat=ios)
      close (13)
      t_lattice_1 = t1; t_lattice_2 = t2; t_lattice_3 = t3; LatticeOrigin = origin
      call get_command_argument(3, fin)
      call get_command_argument(4, fout)
      open (11, file=trim(fin), access='stream', form='unformatted', status='old')
      read (11) n
      allocate (particles(n), a(n))
      do k = 1, 3
         read (11) a
         particles(:)%x(k) = a
      end do
      close (11)
      call constrain_periodic(particles)
      open (12, file=trim(fout), access='stream', form='unformatted', status='replace')
      write (12) n
      do k = 1, 3
         write (12) particles(:)%x(k)
      end do
      close (12)
      call timer_reset_all()
      call timer_start(t_userdefined_first)
      call timer_stop(t_userdefined_first)
      call timer_add(t_userdefined_first + 1, 1.25_8)
      call timings_GatherAndOutput(1)
      call timer_add(t_userdefined_first + 1, 0.5_8)
      call timings_GatherAndOutput(2, 7)
      write (*, '(a,l1)') 'dbg_stats ', dbg(DBG_STATS)
      write (*, '(a,i0)') 'debug_level ', debug_level
      write (*, '(a,3l1)') 'periodicity ', periodicity
      write (*, '(a,i0)') 'mirror_box_layers ', mirror_box_layers
      write (*, '(a,i0)') 'mac_select ', mac_select
      write (*, '(a,es24.16)') 'theta2 ', theta2
      write (*, '(a,es24.16)') 'eps2 ', eps2
      write (*, '(a,es24.16)') 'kelbg_invsqrttemp ', kelbg_invsqrttemp
   end subroutine library_modules

   subroutine vtk_series()
      integer :: k, vstep
      call get_command_argument(2, fin)
      call get_command_argument(3, fout)
      open (11, file=trim(fin), access='stream', form='unformatted', status='old')
      read (11) n
      allocate (particles(n), a(n))
      read (11) a; particles(:)%x(1) = a
      read (11) a; particles(:)%x(2) = a
      read (11) a; particles(:)%x(3) = a
      read (11) a; particles(:)%data%q = a
      close (11)
      do i = 1, n
         particles(i)%label = i - 1
      end do
      do k = 0, 2
         particles(:)%work = 1.0_8 + k
         particles(:)%results%pot = particles(:)%data%q * (k + 1)
         particles(:)%results%e(1) = particles(:)%x(1) * (k + 1)
         particles(:)%results%e(2) = -particles(:)%x(2)
         particles(:)%results%e(3) = particles(:)%x(3) + k
         vstep = VTK_STEP_NORMAL
         if (k == 0) vstep = VTK_STEP_FIRST
         if (k == 2) vstep = VTK_STEP_LAST
         call vtk_write_spacecurve(k, 0.5_8 * k, vstep, particles)
         call vtk_write_branches(k, 0.5_8 * k, vstep, global_tree)
         call vtk_write_leaves(k, 0.5_8 * k, vstep, global_tree)
         select case (k)
         case (1)
            call vtk_write_particles(trim(fout), 0, k, 0.5_8 * k, vstep, particles, with_l2)
         case (2)
            call vtk_write_particles(trim(fout), 0, k, 0.5_8 * k, vstep, particles, coord_scale=2.0_8)
         case default
            call vtk_write_particles(trim(fout), 0, k, 0.5_8 * k, vstep, particles)
         end select
      end do
   end subroutine vtk_series

   subroutine init_particles(p)
      type(t_particle), intent(inout) :: p(:)
      integer(kind_particle) :: j
      integer(8) :: s
      s = 88172645463325252_8
      do j = 1, size(p, kind=kind_particle)
         p(j)%label = j
         p(j)%data%q = -1.0_8 + 2.0_8*mod(j, 2_kind_particle)   ! module_helper.f90:185
         p(j)%x(1) = rnd(s)
         p(j)%x(2) = rnd(s)
         p(j)%x(3) = rnd(s)
         p(j)%work = 1.0_8
      end do
   end subroutine init_particles

   real(8) function rnd(s)
      integer(8), intent(inout) :: s
      ! xorshift64 mapped to [0,1)
      s = ieor(s, ishft(s, 13))
      s = ieor(s, ishft(s, -7))
      s = ieor(s, ishft(s, 17))
      rnd = real(ishft(s, -11), 8)*(1.0_8/9007199254740992.0_8)
   end function rnd

   subroutine push(p, vel, h)
      type(t_particle), intent(inout) :: p(:)
      real(kind_physics), intent(inout) :: vel(:, :)
      real(kind_physics), intent(in) :: h
      integer(kind_particle) :: j, l
      do j = 1, size(p, kind=kind_particle)
         l = p(j)%label
         vel(:, l) = vel(:, l) + h*p(j)%data%q*p(j)%results%e
         p(j)%x = p(j)%x + h*vel(:, l)
      end do
   end subroutine push

end program pepc_mini_gpu
