!
! tlm_gpu_iface.F90 -- iso_c_binding interface of the MI355X SCN engine
! (include/tlmscn.h) for Fortran hosts.
!
! The engine replaces the argument-less hot-path subroutines of GGI_TLM
! v2.1.7 (SRC/TLM_SOLVER/run_TLM.F90:242-283: excitation, scatter,
! cell_output, connect, outer_boundary).  A host keeps its module tables in
! the reference's layout and hands them over once; the thin wrappers below
! turn a nonzero engine status into the reference's error convention
! (write(*,*) message; STOP 1).
!
module tlm_gpu_iface

  use iso_c_binding

  implicit none

  integer(c_int32_t), parameter :: TLM_FORM_PLAIN_FP64 = 0
  integer(c_int32_t), parameter :: TLM_FORM_PRECISION_FP32 = 1
  integer(c_int32_t), parameter :: TLM_FORM_EXTENDED_FP32 = 2
  integer, parameter :: TLM_MAX_FILTER_ORDER = 8

  ! TLM_general grid scalars (tlm_grid)
  type, bind(c) :: tlm_grid
    integer(c_int32_t) :: nx, ny, nz, n_timesteps
    real(c_double) :: dl, dt, reduced_c_factor
    real(c_double) :: r_xmin, r_xmax, r_ymin, r_ymax, r_zmin, r_zmax
  end type tlm_grid

  ! one volume_material_list row (tlm_volume_material)
  type, bind(c) :: tlm_volume_material
    integer(c_int32_t) :: mtype, eps_filter_exists, mu_filter_exists
    integer(c_int32_t) :: eps_order, mu_order, pad_
    real(c_double) :: Ys_eps_f, Ge, Zs_mu_f, Rm
    real(c_double) :: eps_a(0:TLM_MAX_FILTER_ORDER), eps_b(0:TLM_MAX_FILTER_ORDER)
    real(c_double) :: mu_a(0:TLM_MAX_FILTER_ORDER), mu_b(0:TLM_MAX_FILTER_ORDER)
  end type tlm_volume_material

  ! one PML_parameters row without the depth fields (tlm_pml_parameter)
  type, bind(c) :: tlm_pml_parameter
    real(c_double) :: sx, sy, sz, ax, ay, az, exp_x, exp_y, exp_z
    real(c_double) :: Csx, Csy, Csz, Csx2, Csy2, Csz2, Csy_sz, Csz_sx, Csx_sy
  end type tlm_pml_parameter

  ! one point source (tlm_point_source)
  type, bind(c) :: tlm_point_source
    integer(c_int32_t) :: func, component, source_type, at_face, row
  end type tlm_point_source

  ! one Huygens face (tlm_huygens_face)
  type, bind(c) :: tlm_huygens_face
    integer(c_int32_t) :: row, nx, ny, nz
    real(c_double) :: offset
  end type tlm_huygens_face

  ! one frequency output volume or surface (tlm_frequency_output)
  type, bind(c) :: tlm_frequency_output
    real(c_double) :: frequency
    integer(c_int32_t) :: kind, field_component
    integer(c_int32_t) :: first_timestep, last_timestep, timestep_interval
    integer(c_int32_t) :: n_items
  end type tlm_frequency_output

  ! one time-domain surface or volume field output (tlm_field_output)
  type, bind(c) :: tlm_field_output
    integer(c_int32_t) :: kind, field_component, both_sides
    integer(c_int32_t) :: first_timestep, last_timestep, timestep_interval
    integer(c_int32_t) :: n_items
  end type tlm_field_output

  ! one far-field surface (tlm_far_field_surface); angles in radians
  type, bind(c) :: tlm_far_field_surface
    real(c_double) :: frequency
    real(c_double) :: theta_min, theta_step, phi_min, phi_step
    integer(c_int32_t) :: n_theta, n_phi
    integer(c_int32_t) :: dim
    integer(c_int32_t) :: n_faces
  end type tlm_far_field_surface

  ! the RCS surface (tlm_rcs_surface); angles in radians
  type, bind(c) :: tlm_rcs_surface
    real(c_double) :: theta, phi
    integer(c_int32_t) :: n_faces, n_points, point_offset, points_per_cell
  end type tlm_rcs_surface

  interface

    integer(c_int) function tlm_create(grid, form, h) bind(c, name='tlm_create')
      import :: c_int, c_int32_t, c_ptr, tlm_grid
      type(tlm_grid), intent(in) :: grid
      integer(c_int32_t), value :: form
      type(c_ptr), intent(out) :: h
    end function tlm_create

    subroutine tlm_destroy(h) bind(c, name='tlm_destroy')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine tlm_destroy

    type(c_ptr) function tlm_last_error(h) bind(c, name='tlm_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function tlm_last_error

    integer(c_int) function tlm_set_cell_codes(h, code, n_special, to_material, to_pml, to_excitation, &
                                               to_output, to_cable) bind(c, name='tlm_set_cell_codes')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(in) :: code(*)
      integer(c_int32_t), value :: n_special
      integer(c_int32_t), intent(in) :: to_material(*), to_pml(*), to_excitation(*), to_output(*), to_cable(*)
    end function tlm_set_cell_codes

    integer(c_int) function tlm_set_face_codes(h, code, n_special, to_material, to_excitation, to_output, &
                                               to_cable) bind(c, name='tlm_set_face_codes')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(in) :: code(*)
      integer(c_int32_t), value :: n_special
      integer(c_int32_t), intent(in) :: to_material(*), to_excitation(*), to_output(*), to_cable(*)
    end function tlm_set_face_codes

    integer(c_int) function tlm_set_volume_materials(h, n, m) bind(c, name='tlm_set_volume_materials')
      import :: c_int, c_int32_t, c_ptr, tlm_volume_material
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      type(tlm_volume_material), intent(in) :: m(*)
    end function tlm_set_volume_materials

    integer(c_int) function tlm_set_surface_materials(h, n, mtype) bind(c, name='tlm_set_surface_materials')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: mtype(*)
    end function tlm_set_surface_materials

    integer(c_int) function tlm_set_pml(h, n_params, params, n_cells, cell_param) bind(c, name='tlm_set_pml')
      import :: c_int, c_int32_t, c_ptr, tlm_pml_parameter
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n_params, n_cells
      type(tlm_pml_parameter), intent(in) :: params(*)
      integer(c_int32_t), intent(in) :: cell_param(*)
    end function tlm_set_pml

    integer(c_int) function tlm_set_excitation_functions(h, n_functions, value, value_face) &
                                               bind(c, name='tlm_set_excitation_functions')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n_functions
      real(c_double), intent(in) :: value(*), value_face(*)
    end function tlm_set_excitation_functions

    integer(c_int) function tlm_set_point_sources(h, n, sources, n_cell_rows, n_face_rows) &
                                               bind(c, name='tlm_set_point_sources')
      import :: c_int, c_int32_t, c_ptr, tlm_point_source
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n, n_cell_rows, n_face_rows
      type(tlm_point_source), intent(in) :: sources(*)
    end function tlm_set_point_sources

    integer(c_int) function tlm_set_huygens_surface(h, func, offset_min, Ei, Hi, n_faces, faces) &
                                               bind(c, name='tlm_set_huygens_surface')
      import :: c_int, c_int32_t, c_ptr, c_double, tlm_huygens_face
      type(c_ptr), value :: h
      integer(c_int32_t), value :: func, n_faces
      real(c_double), value :: offset_min
      real(c_double), intent(in) :: Ei(3), Hi(3)
      type(tlm_huygens_face), intent(in) :: faces(*)
    end function tlm_set_huygens_surface

    integer(c_int) function tlm_set_output_points(h, n, row, component) bind(c, name='tlm_set_output_points')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: row(*), component(*)
    end function tlm_set_output_points

    integer(c_int) function tlm_set_output_capacity(h, steps) bind(c, name='tlm_set_output_capacity')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: steps
    end function tlm_set_output_capacity

    integer(c_int) function tlm_upload_links(h, V) bind(c, name='tlm_upload_links')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: V(*)
    end function tlm_upload_links

    integer(c_int) function tlm_download_links(h, V) bind(c, name='tlm_download_links')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: V(*)
    end function tlm_download_links

    ! tlm_download_links without device allocation (fixed 32 MiB staging buffer)
    integer(c_int) function tlm_download_links_noalloc(h, V) bind(c, name='tlm_download_links_noalloc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: V(*)
    end function tlm_download_links_noalloc

    integer(c_int) function tlm_excitation(h, timestep) bind(c, name='tlm_excitation')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: timestep
    end function tlm_excitation

    integer(c_int) function tlm_scatter(h, timestep) bind(c, name='tlm_scatter')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: timestep
    end function tlm_scatter

    integer(c_int) function tlm_connect(h, timestep) bind(c, name='tlm_connect')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: timestep
    end function tlm_connect

    integer(c_int) function tlm_outer_boundary(h) bind(c, name='tlm_outer_boundary')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function tlm_outer_boundary

    integer(c_int) function tlm_step(h, first_timestep, n_steps) bind(c, name='tlm_step')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: first_timestep, n_steps
    end function tlm_step

    integer(c_int) function tlm_read_cell_outputs(h, out, max_steps, n_read, first_step) &
                                               bind(c, name='tlm_read_cell_outputs')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: out(*)
      integer(c_int32_t), value :: max_steps
      integer(c_int32_t), intent(out) :: n_read, first_step
    end function tlm_read_cell_outputs

    integer(c_int) function tlm_read_face_outputs(h, out, max_steps, n_read, first_step) &
                                               bind(c, name='tlm_read_face_outputs')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: out(*)
      integer(c_int32_t), value :: max_steps
      integer(c_int32_t), intent(out) :: n_read, first_step
    end function tlm_read_face_outputs

    integer(c_int) function tlm_set_small_to_zero(h, n_small_to_zero, small_to_zero_value) &
                                               bind(c, name='tlm_set_small_to_zero')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n_small_to_zero
      real(c_double), value :: small_to_zero_value
    end function tlm_set_small_to_zero

    integer(c_int) function tlm_set_layout(h, layout) bind(c, name='tlm_set_layout')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: layout
    end function tlm_set_layout

    integer(c_int) function tlm_get_schedule(h, variant, rows, zc, xcd_remap, mixed, streams) &
                                               bind(c, name='tlm_get_schedule')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: variant, rows, zc, xcd_remap, mixed, streams
    end function tlm_get_schedule

    integer(c_int) function tlm_set_schedule(h, variant, rows, zc, xcd_remap, mixed) &
                                               bind(c, name='tlm_set_schedule')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: variant, rows, zc, xcd_remap, mixed
    end function tlm_set_schedule

    integer(c_int) function tlm_synchronize(h) bind(c, name='tlm_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function tlm_synchronize

    integer(c_int64_t) function tlm_device_bytes(h) bind(c, name='tlm_device_bytes')
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: h
    end function tlm_device_bytes

    integer(c_int) function tlm_set_profiling(h, enable) bind(c, name='tlm_set_profiling')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: enable
    end function tlm_set_profiling

    ! launches(TLM_N_KCLASS), ms(TLM_N_KCLASS)
    integer(c_int) function tlm_read_kernel_times(h, launches, ms) bind(c, name='tlm_read_kernel_times')
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), intent(inout) :: launches(*)
      real(c_double), intent(inout) :: ms(*)
    end function tlm_read_kernel_times

    ! cells(TLM_N_KCLASS), bytes(TLM_N_KCLASS)
    integer(c_int) function tlm_kernel_bytes(h, cells, bytes) bind(c, name='tlm_kernel_bytes')
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int64_t), intent(inout) :: cells(*), bytes(*)
    end function tlm_kernel_bytes

    integer(c_int) function tlm_set_streams(h, streams) bind(c, name='tlm_set_streams')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: streams
    end function tlm_set_streams

    integer(c_int) function tlm_layout_segments(h, n_segments, x0) bind(c, name='tlm_layout_segments')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: n_segments
      integer(c_int32_t), intent(out) :: x0(*)
    end function tlm_layout_segments

    ! Host-path (cable) cells: cells = c_null_ptr asks for the count only,
    ! else c_loc of an integer(c_int32_t) array of 3*n_host entries.
    integer(c_int) function tlm_host_cells(h, n_host, cells) bind(c, name='tlm_host_cells')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: n_host
      type(c_ptr), value :: cells
    end function tlm_host_cells

    ! V(1:12, k) = incident voltages of host-path cell k (between scatter and connect)
    integer(c_int) function tlm_gather_host_cells(h, V) bind(c, name='tlm_gather_host_cells')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(out) :: V(12, *)
    end function tlm_gather_host_cells

    integer(c_int) function tlm_scatter_host_cells(h, V) bind(c, name='tlm_scatter_host_cells')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: V(12, *)
    end function tlm_scatter_host_cells

    ! Host-path faces (thin layer, diode, spice, switch): faces = c_null_ptr
    ! asks for the count, else c_loc of 4*n entries (cx, cy, cz, axis).
    integer(c_int) function tlm_host_faces(h, n_faces, faces) bind(c, name='tlm_host_faces')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: n_faces
      type(c_ptr), value :: faces
    end function tlm_host_faces

    ! V(1:4, k): scattered pulses beside host face k (a on the + side cell,
    ! a on the - side cell, b on +, b on -), between scatter and connect
    integer(c_int) function tlm_gather_host_faces(h, V) bind(c, name='tlm_gather_host_faces')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(out) :: V(4, *)
    end function tlm_gather_host_faces

    ! T(1:4, k) = (Va_min, Va_max, Vb_min, Vb_max): the face voltages of the
    ! surface model, e.g. Vy_min, Vy_max, Vz_min, Vz_max of
    ! connect_normal_to_x.F90
    integer(c_int) function tlm_scatter_host_faces(h, T) bind(c, name='tlm_scatter_host_faces')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: T(4, *)
    end function tlm_scatter_host_faces

    integer(c_int) function tlm_set_dft_frequencies(h, n, f) bind(c, name='tlm_set_dft_frequencies')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: f(*)
    end function tlm_set_dft_frequencies

    integer(c_int) function tlm_read_dft(h, re_im) bind(c, name='tlm_read_dft')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: re_im(*)
    end function tlm_read_dft

    integer(c_int) function tlm_set_frequency_outputs(h, n, outputs, rows, sides, normals) &
                                               bind(c, name='tlm_set_frequency_outputs')
      import :: c_int, c_int32_t, c_ptr, tlm_frequency_output
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      type(tlm_frequency_output), intent(in) :: outputs(*)
      integer(c_int32_t), intent(in) :: rows(*), sides(*), normals(*)
    end function tlm_set_frequency_outputs

    integer(c_int) function tlm_read_frequency_output(h, k, re_im) bind(c, name='tlm_read_frequency_output')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k
      real(c_double), intent(inout) :: re_im(*)
    end function tlm_read_frequency_output

    integer(c_int) function tlm_set_field_outputs(h, n, outputs, rows, sides, normals) &
                                               bind(c, name='tlm_set_field_outputs')
      import :: c_int, c_int32_t, c_ptr, tlm_field_output
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      type(tlm_field_output), intent(in) :: outputs(*)
      integer(c_int32_t), intent(in) :: rows(*), sides(*), normals(*)
    end function tlm_set_field_outputs

    ! values(n_items, max_frames), oldest frame first; the frames read are consumed
    integer(c_int) function tlm_read_field_output(h, k, values, max_frames, n_frames, first_timestep) &
                                               bind(c, name='tlm_read_field_output')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k, max_frames
      real(c_double), intent(inout) :: values(*)
      integer(c_int32_t), intent(out) :: n_frames, first_timestep
    end function tlm_read_field_output

    integer(c_int) function tlm_set_far_field_surfaces(h, n, surfaces, rows, sides, normals, centres) &
                                               bind(c, name='tlm_set_far_field_surfaces')
      import :: c_int, c_int32_t, c_ptr, c_double, tlm_far_field_surface
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      type(tlm_far_field_surface), intent(in) :: surfaces(*)
      integer(c_int32_t), intent(in) :: rows(*), sides(*), normals(*)
      real(c_double), intent(in) :: centres(*)
    end function tlm_set_far_field_surfaces

    ! e(4, n_phi * n_theta): Re, Im E_theta, Re, Im E_phi per angle (phi fastest)
    integer(c_int) function tlm_read_far_field(h, k, e) bind(c, name='tlm_read_far_field')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k
      real(c_double), intent(inout) :: e(*)
    end function tlm_read_far_field

    integer(c_int) function tlm_set_rcs_surface(h, surface, rows, sides, normals, centres) &
                                               bind(c, name='tlm_set_rcs_surface')
      import :: c_int, c_int32_t, c_ptr, c_double, tlm_rcs_surface
      type(c_ptr), value :: h
      type(tlm_rcs_surface), intent(in) :: surface
      integer(c_int32_t), intent(in) :: rows(*), sides(*), normals(*)
      real(c_double), intent(in) :: centres(*)
    end function tlm_set_rcs_surface

    ! out(4, n_points): E_theta, E_phi, H_theta, H_phi per far-field point
    integer(c_int) function tlm_read_rcs(h, out) bind(c, name='tlm_read_rcs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: out(*)
    end function tlm_read_rcs

    integer(c_int) function tlm_autotune(h, first_timestep, steps_per_candidate, steps_done) &
                                               bind(c, name='tlm_autotune')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: first_timestep, steps_per_candidate
      integer(c_int32_t), intent(out) :: steps_done
    end function tlm_autotune

  end interface

contains

  ! Reference error convention: write(*,*) message; STOP 1.
  subroutine tlm_check(h, status, what)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (status .eq. 0) return
    text = ''
    if (c_associated(h)) then
      call c_f_pointer(tlm_last_error(h), msg, (/ 512 /))
      do i = 1, 512
        if (msg(i) .eq. c_null_char) exit
        text(i:i) = msg(i)
      end do
    end if
    write(*,*)'ERROR in ', what, ': ', trim(text)
    STOP 1
  end subroutine tlm_check

  ! Drop-in replacements for the reference's per-step calls
  ! (run_TLM.F90:242, 244, 277, 283).
  subroutine gpu_excitation(h, timestep)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: timestep
    call tlm_check(h, tlm_excitation(h, int(timestep, c_int32_t)), 'excitation')
  end subroutine gpu_excitation

  subroutine gpu_scatter(h, timestep)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: timestep
    call tlm_check(h, tlm_scatter(h, int(timestep, c_int32_t)), 'scatter')
  end subroutine gpu_scatter

  subroutine gpu_connect(h, timestep)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: timestep
    call tlm_check(h, tlm_connect(h, int(timestep, c_int32_t)), 'connect')
  end subroutine gpu_connect

  subroutine gpu_outer_boundary(h)
    type(c_ptr), intent(in) :: h
    call tlm_check(h, tlm_outer_boundary(h), 'outer_boundary')
  end subroutine gpu_outer_boundary

end module tlm_gpu_iface
