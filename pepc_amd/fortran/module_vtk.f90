!> VTK XML output of particles for the frontends (utils/module_vtk.f90,
!> utils/module_vtk_helpers.f90): one unstructured-grid piece per process
!> (./vtk/<name>_<step>.<rank>.vtu), the parallel index written by rank 0
!> (./vtk/<name>_<step>.pvtu) and the ParaView / VisIt time series
!> (./<name>.timeseries.pvd, ./<name>.timeseries.visit).  Binary data arrays
!> are base64 text of the big-endian bytes: a 4-byte length block, then the
!> values (byte_order="BigEndian").  This shim runs as one process, so every
!> file set has a single piece.
module module_vtk
   use iso_c_binding
   implicit none
   private

   integer, public, parameter :: VTK_STEP_FIRST = -1    !< opens the time series
   integer, public, parameter :: VTK_STEP_NORMAL = 0
   integer, public, parameter :: VTK_STEP_LAST = 1      !< closes the time series
   integer, public, parameter :: VTK_VERTEX = 1
   integer, public, parameter :: VTK_POLY_LINE = 4
   integer, public, parameter :: VTK_VOXEL = 11

   character(*), parameter :: dir_series = "./"
   character(*), parameter :: dir_pieces = "./vtk/"
   character(*), parameter :: b64 = &
      "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"

   type, public :: vtkfile_unstructured_grid
      character(64) :: stem = ''              !< <name>_<step>
      integer :: fh = -1, fh_par = -1, fh_pvd = -1, fh_visit = -1
      integer :: my_rank = 0, num_pe = 1, vtk_step = VTK_STEP_NORMAL
      real(c_double) :: simtime = 0.0_c_double
      logical, public :: binary = .true.      !< .false.: ascii data arrays
   contains
      procedure :: create => ug_create
      procedure :: create_parallel => ug_create_parallel
      procedure, private :: ug_headers4
      procedure, private :: ug_headers8
      generic :: write_headers => ug_headers4, ug_headers8
      procedure :: startpoints => ug_startpoints
      procedure :: finishpoints => ug_finishpoints
      procedure :: startpointdata => ug_startpointdata
      procedure :: finishpointdata => ug_finishpointdata
      procedure :: startcells => ug_startcells
      procedure :: finishcells => ug_finishcells
      procedure :: startcelldata => ug_startcelldata
      procedure :: finishcelldata => ug_finishcelldata
      procedure :: dont_write_cells => ug_dont_write_cells
      procedure :: write_final => ug_write_final
      procedure :: close => ug_close
      procedure, private :: da_real8
      procedure, private :: da_real8_3
      procedure, private :: da_int8
      procedure, private :: da_int4
      procedure, private :: da_int4_repeat
      procedure, private :: da_int4_scalar
      procedure, private :: da_int8_scalar
      generic :: write_data_array => da_real8, da_real8_3, da_int8, da_int4, da_int4_repeat, &
         da_int4_scalar, da_int8_scalar
   end type vtkfile_unstructured_grid

   interface
      integer(c_int) function vtk_mkdir(path, mode) bind(C, name='mkdir')
         import :: c_int, c_char
         character(kind=c_char), dimension(*) :: path
         integer(c_int), value :: mode
      end function vtk_mkdir
   end interface

contains

   ! ------------------------------------------------------------- encoding
   !> Appends the base64 text of bytes(1:nb) (one block, '=' padded) to the
   !> current record of unit fh.
   subroutine put_base64(fh, bytes, nb)
      integer, intent(in) :: fh
      integer(c_int8_t), intent(in) :: bytes(*)
      integer(c_int64_t), intent(in) :: nb
      character(len=:), allocatable :: out
      integer(c_int64_t) :: i, k, nout
      integer :: b0, b1, b2, g
      nout = 4 * ((nb + 2) / 3)
      allocate (character(len=nout) :: out)
      k = 0
      do i = 1, nb, 3
         b0 = iand(int(bytes(i)), 255)
         b1 = 0; b2 = 0
         if (i + 1 <= nb) b1 = iand(int(bytes(i + 1)), 255)
         if (i + 2 <= nb) b2 = iand(int(bytes(i + 2)), 255)
         g = ior(ior(ishft(b0, 16), ishft(b1, 8)), b2)
         out(k + 1:k + 1) = b64(ishft(g, -18) + 1:ishft(g, -18) + 1)
         out(k + 2:k + 2) = b64(iand(ishft(g, -12), 63) + 1:iand(ishft(g, -12), 63) + 1)
         if (i + 1 <= nb) then
            out(k + 3:k + 3) = b64(iand(ishft(g, -6), 63) + 1:iand(ishft(g, -6), 63) + 1)
         else
            out(k + 3:k + 3) = '='
         end if
         if (i + 2 <= nb) then
            out(k + 4:k + 4) = b64(iand(g, 63) + 1:iand(g, 63) + 1)
         else
            out(k + 4:k + 4) = '='
         end if
         k = k + 4
      end do
      write (fh, '(a)', advance='no') out
   end subroutine put_base64

   !> A value 0..255 as a (two's-complement) byte.
   elemental integer(c_int8_t) function to_byte(x)
      integer(c_int64_t), intent(in) :: x
      to_byte = int(merge(x - 256_c_int64_t, x, x > 127_c_int64_t), c_int8_t)
   end function to_byte

   !> Big-endian bytes of n 8-byte (be_bytes8) or 4-byte (be_bytes4) integers.
   subroutine be_bytes8(v, n, bytes)
      integer(c_int64_t), intent(in) :: v(:)
      integer(c_int64_t), intent(in) :: n
      integer(c_int8_t), intent(out) :: bytes(:)
      integer(c_int64_t) :: i
      integer :: j
      do i = 1, n
         do j = 0, 7
            bytes(8 * (i - 1) + j + 1) = to_byte(iand(ishft(v(i), -8 * (7 - j)), 255_c_int64_t))
         end do
      end do
   end subroutine be_bytes8

   subroutine be_bytes4(v, n, bytes)
      integer(c_int32_t), intent(in) :: v(:)
      integer(c_int64_t), intent(in) :: n
      integer(c_int8_t), intent(out) :: bytes(:)
      integer(c_int64_t) :: i
      integer :: j
      do i = 1, n
         do j = 0, 3
            bytes(4 * (i - 1) + j + 1) = to_byte(int(iand(ishft(v(i), -8 * (3 - j)), 255_c_int32_t), c_int64_t))
         end do
      end do
   end subroutine be_bytes4

   !> Length block then value block, both base64, on the current record.
   subroutine put_binary8(fh, v)
      integer, intent(in) :: fh
      integer(c_int64_t), intent(in) :: v(:)
      integer(c_int8_t) :: hdr(4)
      integer(c_int8_t), allocatable :: bytes(:)
      integer(c_int64_t) :: n
      n = size(v, kind=c_int64_t)
      call be_bytes4([int(8 * n, c_int32_t)], 1_c_int64_t, hdr)
      call put_base64(fh, hdr, 4_c_int64_t)
      allocate (bytes(8 * max(n, 1_c_int64_t)))
      call be_bytes8(v, n, bytes)
      call put_base64(fh, bytes, 8 * n)
   end subroutine put_binary8

   subroutine put_binary4(fh, v)
      integer, intent(in) :: fh
      integer(c_int32_t), intent(in) :: v(:)
      integer(c_int8_t) :: hdr(4)
      integer(c_int8_t), allocatable :: bytes(:)
      integer(c_int64_t) :: n
      n = size(v, kind=c_int64_t)
      call be_bytes4([int(4 * n, c_int32_t)], 1_c_int64_t, hdr)
      call put_base64(fh, hdr, 4_c_int64_t)
      allocate (bytes(4 * max(n, 1_c_int64_t)))
      call be_bytes4(v, n, bytes)
      call put_base64(fh, bytes, 4 * n)
   end subroutine put_binary4

   ! ------------------------------------------------------------- files
   subroutine ug_create(vtk, filename_, step_, simtime_, vtk_step_)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: filename_
      integer :: step_, vtk_step_
      real(c_double) :: simtime_
      call vtk%create_parallel(filename_, step_, 0, 1, simtime_, vtk_step_)
   end subroutine ug_create

   subroutine ug_create_parallel(vtk, filename_, step_, my_rank_, num_pe_, simtime_, vtk_step_, comm_)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: filename_
      integer :: step_, my_rank_, num_pe_, vtk_step_
      real(c_double) :: simtime_
      integer, optional :: comm_
      character(6) :: rank
      integer(c_int) :: rc
      vtk%my_rank = my_rank_
      vtk%num_pe = max(num_pe_, 1)
      vtk%simtime = simtime_
      vtk%vtk_step = vtk_step_
      write (vtk%stem, '(a, "_", i6.6)') filename_, step_
      write (rank, '(i6.6)') vtk%my_rank
      rc = vtk_mkdir(dir_pieces//c_null_char, int(o'755', c_int))
      open (newunit=vtk%fh, file=dir_pieces//trim(vtk%stem)//"."//rank//".vtu", status='replace')
      if (vtk%my_rank == 0) then
         open (newunit=vtk%fh_par, file=dir_pieces//trim(vtk%stem)//".pvtu", status='replace')
         if (vtk%vtk_step == VTK_STEP_FIRST) then
            open (newunit=vtk%fh_visit, file=dir_series//filename_//".timeseries.visit", status='replace')
            write (vtk%fh_visit, '("!NBLOCKS ", i0)') vtk%num_pe
            open (newunit=vtk%fh_pvd, file=dir_series//filename_//".timeseries.pvd", status='replace')
            write (vtk%fh_pvd, '(a)') '<VTKFile type="Collection">'
            write (vtk%fh_pvd, '(a)') '<Collection>'
         else
            open (newunit=vtk%fh_visit, file=dir_series//filename_//".timeseries.visit", &
                  status='unknown', position='append')
            open (newunit=vtk%fh_pvd, file=dir_series//filename_//".timeseries.pvd", &
                  status='unknown', position='append')
         end if
      end if
   end subroutine ug_create_parallel

   subroutine both(vtk, piece, par)
      class(vtkfile_unstructured_grid) :: vtk
      character(*), intent(in) :: piece, par
      write (vtk%fh, '(a)') piece
      if (vtk%my_rank == 0) write (vtk%fh_par, '(a)') par
   end subroutine both

   subroutine ug_headers8(vtk, npart, ncell)
      class(vtkfile_unstructured_grid) :: vtk
      integer(c_int64_t) :: npart, ncell
      write (vtk%fh, '(a)') '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="BigEndian">'
      write (vtk%fh, '(a)') '<UnstructuredGrid GhostLevel="0">'
      write (vtk%fh, '("<Piece NumberOfPoints=""", i0, """ NumberOfCells=""", i0, """>")') npart, ncell
      if (vtk%my_rank == 0) then
         write (vtk%fh_par, '(a)') '<VTKFile type="PUnstructuredGrid" version="0.1" byte_order="BigEndian">'
         write (vtk%fh_par, '(a)') '<PUnstructuredGrid GhostLevel="0">'
      end if
   end subroutine ug_headers8

   subroutine ug_headers4(vtk, npart, ncell)
      class(vtkfile_unstructured_grid) :: vtk
      integer :: npart, ncell
      call ug_headers8(vtk, int(npart, c_int64_t), int(ncell, c_int64_t))
   end subroutine ug_headers4

   subroutine ug_startpoints(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call both(vtk, '<Points>', '<PPoints>')
   end subroutine ug_startpoints

   subroutine ug_finishpoints(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call both(vtk, '</Points>', '</PPoints>')
   end subroutine ug_finishpoints

   subroutine ug_startpointdata(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call both(vtk, '<PointData>', '<PPointData>')
   end subroutine ug_startpointdata

   subroutine ug_finishpointdata(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call both(vtk, '</PointData>', '</PPointData>')
   end subroutine ug_finishpointdata

   subroutine ug_startcells(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call both(vtk, '<Cells>', '<PCells>')
   end subroutine ug_startcells

   subroutine ug_finishcells(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call both(vtk, '</Cells>', '</PCells>')
   end subroutine ug_finishcells

   subroutine ug_startcelldata(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call both(vtk, '<CellData>', '<PCellData>')
   end subroutine ug_startcelldata

   subroutine ug_finishcelldata(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call both(vtk, '</CellData>', '</PCellData>')
   end subroutine ug_finishcelldata

   character(6) function fmt_name(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      fmt_name = merge('binary', 'ascii ', vtk%binary)
   end function fmt_name

   subroutine ug_dont_write_cells(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      call vtk%startcells()
      write (vtk%fh, '(3a)') '<DataArray type="Int32" Name="connectivity" format="', trim(fmt_name(vtk)), '" />'
      write (vtk%fh, '(3a)') '<DataArray type="Int32" Name="offsets" format="', trim(fmt_name(vtk)), '" />'
      write (vtk%fh, '(3a)') '<DataArray type="UInt8" Name="types" format="', trim(fmt_name(vtk)), '" />'
      call vtk%finishcells()
      call vtk%startcelldata()
      call vtk%finishcelldata()
   end subroutine ug_dont_write_cells

   subroutine ug_write_final(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      character(6) :: rank
      character(32) :: tbuf
      integer :: i
      write (vtk%fh, '(a)') '</Piece>'
      write (vtk%fh, '(a)') '</UnstructuredGrid>'
      write (vtk%fh, '(a)') '</VTKFile>'
      if (vtk%my_rank == 0) then
         write (vtk%fh_visit, '(a)') ''
         do i = 0, vtk%num_pe - 1
            write (rank, '(i6.6)') i
            write (vtk%fh_par, '(3a)') '<Piece Source="', trim(vtk%stem)//"."//rank//".vtu", '"/>'
            write (vtk%fh_visit, '(a)') dir_pieces//trim(vtk%stem)//"."//rank//".vtu"
         end do
         write (vtk%fh_par, '(a)') '</PUnstructuredGrid>'
         write (vtk%fh_par, '(a)') '</VTKFile>'
         write (tbuf, '(es23.15)') vtk%simtime
         write (vtk%fh_pvd, '(5a)') '<DataSet timestep="', trim(adjustl(tbuf)), '" file="', &
            dir_pieces//trim(vtk%stem)//".pvtu", '"/>'
      end if
   end subroutine ug_write_final

   subroutine ug_close(vtk)
      class(vtkfile_unstructured_grid) :: vtk
      close (vtk%fh)
      if (vtk%my_rank == 0) then
         if (vtk%vtk_step == VTK_STEP_LAST) then
            write (vtk%fh_pvd, '(a)') '</Collection>'
            write (vtk%fh_pvd, '(a)') '</VTKFile>'
         end if
         close (vtk%fh_par)
         close (vtk%fh_visit)
         close (vtk%fh_pvd)
      end if
   end subroutine ug_close

   ! ------------------------------------------------------------- data arrays
   subroutine da_header(vtk, name, ncomp, vtype)
      class(vtkfile_unstructured_grid) :: vtk
      character(*), intent(in) :: name, vtype
      integer, intent(in) :: ncomp
      write (vtk%fh, '(3a, i0, 5a)') '<DataArray Name="', name, '" NumberOfComponents="', ncomp, &
         '" type="', vtype, '" format="', trim(fmt_name(vtk)), '">'
      if (vtk%my_rank == 0) write (vtk%fh_par, '(3a, i0, 5a)') '<DataArray Name="', name, &
         '" NumberOfComponents="', ncomp, '" type="', vtype, '" format="', trim(fmt_name(vtk)), '"/>'
   end subroutine da_header

   subroutine da_real8(vtk, name, data, scale)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: name
      real(c_double) :: data(:)
      real(c_double), optional, intent(in) :: scale
      real(c_double), allocatable :: v(:)
      call da_header(vtk, name, 1, "Float64")
      v = data
      if (present(scale)) v = v * scale
      call da_values8(vtk, transfer(v, [0_c_int64_t], size(v)), v)
   end subroutine da_real8

   subroutine da_real8_3(vtk, name, data1, data2, data3, scale)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: name
      real(c_double) :: data1(:), data2(:), data3(:)
      real(c_double), optional, intent(in) :: scale
      real(c_double), allocatable :: v(:)
      integer(c_int64_t) :: i, n
      call da_header(vtk, name, 3, "Float64")
      n = size(data1, kind=c_int64_t)
      allocate (v(3 * n))
      do i = 1, n
         v(3 * i - 2) = data1(i); v(3 * i - 1) = data2(i); v(3 * i) = data3(i)
      end do
      if (present(scale)) v = scale * v
      call da_values8(vtk, transfer(v, [0_c_int64_t], size(v)), v)
   end subroutine da_real8_3

   !> Writes 8-byte values: bit patterns w (binary) or the reals r (ascii).
   subroutine da_values8(vtk, w, r)
      class(vtkfile_unstructured_grid) :: vtk
      integer(c_int64_t), intent(in) :: w(:)
      real(c_double), intent(in) :: r(:)
      integer(c_int64_t) :: i
      if (vtk%binary) then
         call put_binary8(vtk%fh, w)
      else
         do i = 1, size(r, kind=c_int64_t)
            write (vtk%fh, '(es24.16)') r(i)
         end do
      end if
      write (vtk%fh, '(a)') '</DataArray>'
   end subroutine da_values8

   subroutine da_int8(vtk, name, data)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: name
      integer(c_int64_t) :: data(:)
      integer(c_int64_t) :: i
      call da_header(vtk, name, 1, "Int64")
      if (vtk%binary) then
         call put_binary8(vtk%fh, data)
      else
         do i = 1, size(data, kind=c_int64_t)
            write (vtk%fh, '(i20)') data(i)
         end do
      end if
      write (vtk%fh, '(a)') '</DataArray>'
   end subroutine da_int8

   subroutine da_int4(vtk, name, data)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: name
      integer(c_int32_t) :: data(:)
      integer(c_int64_t) :: i
      call da_header(vtk, name, 1, "Int32")
      if (vtk%binary) then
         call put_binary4(vtk%fh, data)
      else
         do i = 1, size(data, kind=c_int64_t)
            write (vtk%fh, '(i12)') data(i)
         end do
      end if
      write (vtk%fh, '(a)') '</DataArray>'
   end subroutine da_int4

   !> A one-element array (e.g. the offset or type of a single cell).
   subroutine da_int4_scalar(vtk, name, value)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: name
      integer(c_int32_t) :: value
      call da_int4(vtk, name, [value])
   end subroutine da_int4_scalar

   subroutine da_int8_scalar(vtk, name, value)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: name
      integer(c_int64_t) :: value
      call da_int8(vtk, name, [value])
   end subroutine da_int8_scalar

   !> n copies of one value (e.g. the writing rank per point).
   subroutine da_int4_repeat(vtk, name, n, value)
      class(vtkfile_unstructured_grid) :: vtk
      character(*) :: name
      integer(c_int32_t) :: n, value
      integer(c_int32_t), allocatable :: v(:)
      allocate (v(n))
      v = value
      call da_int4(vtk, name, v)
   end subroutine da_int4_repeat
end module module_vtk

!> Particle and tree output of the frontends (utils/module_vtk_helpers.f90:49-113, 233-723,
!> coulomb/module_interaction_specific_types_XYZQ.f90:63-75).
module module_vtk_helpers
   use iso_c_binding
   use module_vtk
   use module_pepc_types
   use module_interaction_specific_types
   use module_spacefilling, only: key_to_coord
   use module_pepc, only: pepc_device_tree
   implicit none
   private
   public :: vtk_write_particles, vtk_write_particle_data_results, vtk_write_spacecurve
   public :: vtk_write_branches, vtk_write_leaves, vtk_write_node_interaction_data

contains

   !> Positions, work, label, local index and rank per particle, the
   !> caller's data arrays (helper_func; default: charge, potential, field)
   !> and one vertex cell per particle.  The communicator is accepted for
   !> the frontends' call and unused: this shim writes as rank 0 of 1.
   subroutine vtk_write_particles(fname, mpi_communicator, step, time, vtk_step, p, helper_func, coord_scale)
      character(*), intent(in) :: fname
      integer, intent(in) :: mpi_communicator
      integer, intent(in) :: step
      real*8, intent(in) :: time
      integer, intent(in) :: vtk_step
      type(t_particle), intent(in) :: p(:)
      interface
         subroutine helper_func(d, r, vtkf)
            use module_interaction_specific_types, only: t_particle_data, t_particle_results
            use module_vtk, only: vtkfile_unstructured_grid
            implicit none
            type(t_particle_data), intent(in) :: d(:)
            type(t_particle_results), intent(in) :: r(:)
            type(vtkfile_unstructured_grid), intent(inout) :: vtkf
         end subroutine helper_func
      end interface
      optional :: helper_func
      real(kind_physics), intent(in), optional :: coord_scale
      type(vtkfile_unstructured_grid) :: vtk
      integer(kind_particle) :: i, np

      np = size(p, kind=kind_particle)
      call vtk%create_parallel(fname, step, 0, 1, time, vtk_step)
      call vtk%write_headers(np, np)
      call vtk%startpoints()
      call vtk%write_data_array("xyz", p(:)%x(1), p(:)%x(2), p(:)%x(3), coord_scale)
      call vtk%finishpoints()
      call vtk%startpointdata()
      call vtk%write_data_array("work", p(:)%work)
      call vtk%write_data_array("label", p(:)%label)
      call vtk%write_data_array("local index", [(i, i=1, np)])
      call vtk%write_data_array("processor", int(np, c_int32_t), 0_c_int32_t)
      if (present(helper_func)) then
         call helper_func(p(:)%data, p(:)%results, vtk)
      else
         call vtk_write_particle_data_results(p(:)%data, p(:)%results, vtk)
      end if
      call vtk%finishpointdata()
      call vtk%startcells()
      call vtk%write_data_array("connectivity", [(i, i=0, np - 1)])
      call vtk%write_data_array("offsets", [(i, i=1, np)])
      call vtk%write_data_array("types", [(int(VTK_VERTEX, c_int32_t), i=1, np)])
      call vtk%finishcells()
      call vtk%startcelldata()
      call vtk%finishcelldata()
      call vtk%write_final()
      call vtk%close()
   end subroutine vtk_write_particles

   !> The space-filling curve: the particles, in the order given (the key
   !> order after a grow), joined by one poly-line cell
   !> (utils/module_vtk_helpers.f90:724-764).
   subroutine vtk_write_spacecurve(step, tsim, vtk_step, particles, coord_scale)
      integer, intent(in) :: step
      integer, intent(in) :: vtk_step
      real*8, intent(in) :: tsim
      type(t_particle), intent(in) :: particles(:)
      real(kind_physics), intent(in), optional :: coord_scale
      type(vtkfile_unstructured_grid) :: vtk
      integer(kind_particle) :: i, npp
      npp = size(particles, kind=kind_particle)
      call vtk%create_parallel("spacecurve", step, 0, 1, tsim, vtk_step)
      call vtk%write_headers(npp, 1_kind_particle)
      call vtk%startpoints()
      call vtk%write_data_array("xyz", particles(:)%x(1), particles(:)%x(2), particles(:)%x(3), coord_scale)
      call vtk%finishpoints()
      call vtk%startpointdata()
      call vtk%finishpointdata()
      call vtk%startcells()
      call vtk%write_data_array("connectivity", [(i, i=0, npp - 1)])
      call vtk%write_data_array("offsets", npp)
      call vtk%write_data_array("types", int(VTK_POLY_LINE, c_int32_t))
      call vtk%finishcells()
      call vtk%startcelldata()
      call vtk%write_data_array("processor", 0_c_int32_t)
      call vtk%finishcelldata()
      call vtk%write_final()
      call vtk%close()
   end subroutine vtk_write_spacecurve

   !> Leaves of the grown tree as boxes (utils/module_vtk_helpers.f90:553-620):
   !> one voxel per leaf, in key order (the order of the reference's
   !> depth-first collection), from the device tree.
   subroutine vtk_write_leaves(step, tsim, vtk_step, t, helper_func, coord_scale)
      integer, intent(in) :: step
      integer, intent(in) :: vtk_step
      real*8, intent(in) :: tsim
      type(t_tree), intent(in) :: t
      interface
         subroutine helper_func(d, vtkf)
            use module_interaction_specific_types, only: t_tree_node_interaction_data
            use module_vtk, only: vtkfile_unstructured_grid
            implicit none
            type(t_tree_node_interaction_data), intent(in) :: d(:)
            type(vtkfile_unstructured_grid), intent(inout) :: vtkf
         end subroutine helper_func
      end interface
      optional :: helper_func
      real(kind_physics), intent(in), optional :: coord_scale
      integer(c_int32_t), allocatable :: link(:, :), first(:)
      real(c_double), allocatable :: mom(:, :)
      integer(kind_key), allocatable :: nkey(:)
      integer(kind_node), allocatable :: sel(:)
      integer(kind_node) :: i
      if (.not. t%allocated) then
         write (*, *) 'vtk_write_leaves(): tree is not allocated, aborting leaves output.'
         return
      end if
      call pepc_device_tree(link, first, mom, nkey)
      allocate (sel(t%nleaf_me))
      do i = 1, size(first, kind=kind_node)
         if (link(2, i) == 0) sel(first(i) + 1) = i     ! one leaf per particle, key order
      end do
      call write_nodes_as_boxes("leaves", step, tsim, vtk_step, t, link, mom, nkey, sel, &
                                helper_func, coord_scale)
   end subroutine vtk_write_leaves

   !> Branch nodes of the grown tree as boxes (utils/module_vtk_helpers.f90:
   !> 625-700).  On one rank the branches are the root's children
   !> (find_local_branches, module_tree_grow.f90:283-363, with the whole key
   !> range as the rank's domain), in key order.
   subroutine vtk_write_branches(step, tsim, vtk_step, t, helper_func, coord_scale)
      integer, intent(in) :: step
      integer, intent(in) :: vtk_step
      real*8, intent(in) :: tsim
      type(t_tree), intent(in) :: t
      interface
         subroutine helper_func(d, vtkf)
            use module_interaction_specific_types, only: t_tree_node_interaction_data
            use module_vtk, only: vtkfile_unstructured_grid
            implicit none
            type(t_tree_node_interaction_data), intent(in) :: d(:)
            type(vtkfile_unstructured_grid), intent(inout) :: vtkf
         end subroutine helper_func
      end interface
      optional :: helper_func
      real(kind_physics), intent(in), optional :: coord_scale
      integer(c_int32_t), allocatable :: link(:, :), first(:)
      real(c_double), allocatable :: mom(:, :)
      integer(kind_key), allocatable :: nkey(:)
      integer(kind_node), allocatable :: sel(:)
      integer(kind_node) :: i, j, nb, tmp
      if (.not. t%allocated) then
         write (*, *) 'vtk_write_branches(): tree is not allocated, aborting branch output.'
         return
      end if
      call pepc_device_tree(link, first, mom, nkey)
      allocate (sel(t%nbranch))
      nb = 0
      do i = 1, size(first, kind=kind_node)
         if (link(4, i) == 1) then
            nb = nb + 1
            sel(nb) = i
         end if
      end do
      do i = 2, nb                                     ! key order (at most eight)
         j = i
         do while (j > 1)
            if (nkey(sel(j - 1)) <= nkey(sel(j))) exit
            tmp = sel(j); sel(j) = sel(j - 1); sel(j - 1) = tmp
            j = j - 1
         end do
      end do
      call write_nodes_as_boxes("branches", step, tsim, vtk_step, t, link, mom, nkey, sel(1:nb), &
                                helper_func, coord_scale)
   end subroutine vtk_write_branches

   !> vtk_write_nodes_as_boxes (utils/module_vtk_helpers.f90:233-373) for the
   !> nodes sel of the device tree: eight corners per node (the cell's
   !> key_to_coord point shifted by the node's box size along the voxel's
   !> axes), voxel cells, and per node the rank, level, leaves, descendants
   !> and the interaction data (helper_func, or charge and abs_charge).
   subroutine write_nodes_as_boxes(fname, step, tsim, vtk_step, t, link, mom, nkey, sel, &
                                   helper_func, coord_scale)
      character(*), intent(in) :: fname
      integer, intent(in) :: step, vtk_step
      real*8, intent(in) :: tsim
      type(t_tree), intent(in) :: t
      integer(c_int32_t), intent(in) :: link(:, :)
      real(c_double), intent(in) :: mom(:, :)
      integer(kind_key), intent(in) :: nkey(:)
      integer(kind_node), intent(in) :: sel(:)
      interface
         subroutine helper_func(d, vtkf)
            use module_interaction_specific_types, only: t_tree_node_interaction_data
            use module_vtk, only: vtkfile_unstructured_grid
            implicit none
            type(t_tree_node_interaction_data), intent(in) :: d(:)
            type(vtkfile_unstructured_grid), intent(inout) :: vtkf
         end subroutine helper_func
      end interface
      optional :: helper_func
      real(kind_physics), intent(in), optional :: coord_scale
      ! corner j of a voxel: offsets (x, y, z) in units of the box size
      real, parameter :: box_shift(3, 8) = reshape([0., 0., 0., 0., 0., 1., 0., 1., 0., 0., 1., 1., &
                                                    1., 0., 0., 1., 0., 1., 1., 1., 0., 1., 1., 1.], [3, 8])
      real(kind_physics), allocatable :: cx(:), cy(:), cz(:)
      integer(kind_node), allocatable :: cidx(:), coff(:), desc(:), leaves(:)
      integer(c_int32_t), allocatable :: level(:), owner(:)
      type(t_tree_node_interaction_data), allocatable :: d(:)
      type(vtkfile_unstructured_grid) :: vtk
      real(kind_physics) :: bx(3), bsize(3), bshift(3)
      integer(kind_node) :: i, nn, nsel, k, c
      integer :: j

      ! descendants of every node: children plus their descendants
      ! (tree_node_update_from_children, module_tree_grow.f90:791, 817);
      ! a twig's children follow it in the device order
      nn = size(link, 2, kind=kind_node)
      allocate (desc(nn))
      desc = 0
      do i = nn, 1, -1
         if (link(2, i) > 0) then
            do k = 0, link(2, i) - 1
               c = int(link(1, i), kind_node) + k + 1
               desc(i) = desc(i) + 1 + desc(c)
            end do
         end if
      end do

      nsel = size(sel, kind=kind_node)
      allocate (cx(8 * nsel), cy(8 * nsel), cz(8 * nsel), cidx(8 * nsel), coff(nsel), &
                leaves(nsel), level(nsel), owner(nsel), d(nsel))
      do i = 1, nsel
         k = sel(i)
         level(i) = link(4, k)
         leaves(i) = link(3, k)
         owner(i) = 0
         d(i)%coc = mom(1:3, k)
         d(i)%charge = mom(4, k)
         d(i)%abs_charge = mom(5, k)
         d(i)%dip = mom(6:8, k)
         d(i)%quad = mom(9:11, k)
         d(i)%xyquad = mom(12, k)
         d(i)%yzquad = mom(13, k)
         d(i)%zxquad = mom(14, k)
         d(i)%bmax = mom(15, k)
         bsize = t%bounding_box%boxsize / 2**level(i)
         coff(i) = 8 * i
         call key_to_coord(t%bounding_box, nkey(k), bx)
         do j = 1, 8
            cidx(8 * (i - 1) + j) = 8 * (i - 1) + j - 1
            bshift = box_shift(:, j) * bsize
            cx(8 * (i - 1) + j) = bx(1) + bshift(1)
            cy(8 * (i - 1) + j) = bx(2) + bshift(2)
            cz(8 * (i - 1) + j) = bx(3) + bshift(3)
         end do
      end do

      call vtk%create_parallel(fname, step, 0, 1, tsim, vtk_step)
      call vtk%write_headers(8 * nsel, nsel)
      call vtk%startpoints()
      call vtk%write_data_array("corners", cx, cy, cz, coord_scale)
      call vtk%finishpoints()
      call vtk%startpointdata()
      call vtk%finishpointdata()
      call vtk%startcells()
      call vtk%write_data_array("connectivity", cidx)
      call vtk%write_data_array("offsets", coff)
      call vtk%write_data_array("types", int(nsel, c_int32_t), int(VTK_VOXEL, c_int32_t))
      call vtk%finishcells()
      call vtk%startcelldata()
      call vtk%write_data_array("processor", owner)
      call vtk%write_data_array("level", level)
      call vtk%write_data_array("leaves", leaves)
      call vtk%write_data_array("descendants", [(desc(sel(i)), i=1, nsel)])
      if (present(helper_func)) then
         call helper_func(d, vtk)
      else
         call vtk_write_node_interaction_data(d, vtk)
      end if
      call vtk%finishcelldata()
      call vtk%write_final()
      call vtk%close()
   end subroutine write_nodes_as_boxes

   !> coulomb/module_interaction_specific_types_XYZQ.f90:80-89
   subroutine vtk_write_node_interaction_data(d, vtkf)
      type(t_tree_node_interaction_data), intent(in) :: d(:)
      type(vtkfile_unstructured_grid), intent(inout) :: vtkf
      call vtkf%write_data_array("charge", d(:)%charge)
      call vtkf%write_data_array("abs_charge", d(:)%abs_charge)
   end subroutine vtk_write_node_interaction_data

   subroutine vtk_write_particle_data_results(d, r, vtkf)
      type(t_particle_data), intent(in) :: d(:)
      type(t_particle_results), intent(in) :: r(:)
      type(vtkfile_unstructured_grid), intent(inout) :: vtkf
      call vtkf%write_data_array("q", d(:)%q)
      call vtkf%write_data_array("pot", r(:)%pot)
      call vtkf%write_data_array("field", r(:)%e(1), r(:)%e(2), r(:)%e(3))
   end subroutine vtk_write_particle_data_results
end module module_vtk_helpers
