!
! tlm_gpu_cavity.F90 -- Fortran host driving the MI355X engine through
! tlm_gpu_iface, the way run_TLM would with the one-flag switch on.
!
! It builds, in Fortran and in the reference's own table layout, the
! CAVITY_FREE_SPACE deck (TEST_DATA/CAVITY_FREE_SPACE: 10x20x30 PEC cavity,
! soft Ex Gaussian source at cell (3,6,17), Ex probe at cell (7,3,14)),
! runs the reference's per-step call sequence (run_TLM.F90:240-283) on the
! GPU, and writes the probe file in the reference's .tout format
! (write_time_domain_header_data, FILE_HANDLING/read_write_header_data.F90;
! format (E16.7,I5,E16.7), File_information.F90:235; |v| < 1e-30 printed as 0,
! OUTPUT/output_points.F90:597).
!
! Usage: tlm_gpu_cavity <out.tout> [plain|precision]
!
program tlm_gpu_cavity

  use iso_c_binding
  use tlm_gpu_iface

  implicit none

  real(c_double), parameter :: c0 = 2.99792458d8
  integer, parameter :: nx = 10, ny = 20, nz = 30
  real(c_double) :: mesh_min(3), mesh_max(3), dlx, dly, dlz, dl, dt, time, value
  integer :: n_timesteps, timestep, cx, cy, cz, cell, n_faces, n_special
  integer(c_int32_t), allocatable :: cell_code(:), face_code(:)
  integer(c_int32_t) :: to_material(2, 2), zeros2(2)
  integer(c_int32_t), allocatable :: to_pml(:), to_exc(:), to_out(:), to_cable(:)
  integer(c_int32_t) :: dummy(1), row(1), comp(1), n_read, first_step
  real(c_double), allocatable :: fvalue(:), fvalue_face(:), rows(:)
  type(tlm_grid) :: grid
  type(tlm_point_source) :: src(1)
  type(tlm_volume_material) :: vm(1)
  type(c_ptr) :: h
  real(c_double) :: amplitude, width, delay, time2
  character(len=256) :: outname, mode
  integer(c_int32_t) :: form
  integer :: src_cell, out_cell

  call get_command_argument(1, outname)
  mode = 'plain'
  if (command_argument_count() .ge. 2) call get_command_argument(2, mode)
  form = TLM_FORM_PLAIN_FP64
  if (trim(mode) .eq. 'precision') form = TLM_FORM_PRECISION_FP32
  if (trim(mode) .eq. 'extended') form = TLM_FORM_EXTENDED_FP32

  ! mesh (set_mesh_parameters.F90:62-69) and time step (set_TLM_solver_parameters.F90:94-96)
  mesh_min = (/ -0.5d0, -1.0d0, -1.5d0 /)
  mesh_max = (/ 0.5d0, 1.0d0, 1.5d0 /)
  dlx = (mesh_max(1) - mesh_min(1)) / dble(nx)
  dly = (mesh_max(2) - mesh_min(2)) / dble(ny)
  dlz = (mesh_max(3) - mesh_min(3)) / dble(nz)
  dl = (dlx + dly + dlz) / (3d0)
  dt = dl / (2d0 * c0 / 1d0)
  n_timesteps = NINT(1d-6 / dt)

  grid%nx = nx; grid%ny = ny; grid%nz = nz; grid%n_timesteps = n_timesteps
  grid%dl = dl; grid%dt = dt; grid%reduced_c_factor = 1d0
  grid%r_xmin = -1d0; grid%r_xmax = -1d0; grid%r_ymin = -1d0
  grid%r_ymax = -1d0; grid%r_zmin = -1d0; grid%r_zmax = -1d0

  ! cell codes in scan order (set_update_codes.F90:38-363): the probe cell
  ! (7,3,14) precedes the source cell (3,6,17)
  allocate (cell_code(nx * ny * nz))
  cell_code = 0
  out_cell = 7 + nx * ((3 - 1) + ny * (14 - 1))
  src_cell = 3 + nx * ((6 - 1) + ny * (17 - 1))
  n_special = 0
  cell = 0
  allocate (to_pml(2), to_exc(2), to_out(2), to_cable(2))
  to_pml = 0; to_exc = 0; to_out = 0; to_cable = 0
  do cz = 1, nz
    do cy = 1, ny
      do cx = 1, nx
        cell = cell + 1
        if (cell .eq. out_cell .or. cell .eq. src_cell) then
          n_special = n_special + 1
          cell_code(cell) = n_special
          if (cell .eq. out_cell) to_out(n_special) = 1
          if (cell .eq. src_cell) to_exc(n_special) = 1
        end if
      end do
    end do
  end do
  n_faces = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
  allocate (face_code(n_faces))
  face_code = 0
  to_material = 0
  zeros2 = 0
  dummy = 0

  ! Gaussian excitation table (excitation.F90:218-238)
  allocate (fvalue(n_timesteps), fvalue_face(n_timesteps))
  amplitude = 1d0; width = 0.25d-9; delay = 1d-9
  do timestep = 1, n_timesteps
    time = (timestep - 1) * dt
    time2 = time + dt / 2d0
    fvalue(timestep) = 0d0
    fvalue_face(timestep) = 0d0
    if (abs(time - delay) .lt. width * 5d0) fvalue(timestep) = amplitude * (exp(-((time - delay) / width)**2))
    if (abs(time2 - delay) .lt. width * 5) fvalue_face(timestep) = amplitude * (exp(-((time2 - delay) / width)**2))
  end do
  src(1)%func = 1; src(1)%component = 1; src(1)%source_type = 1; src(1)%at_face = 0; src(1)%row = 1
  row(1) = 1; comp(1) = 1

  call tlm_check(c_null_ptr, tlm_create(grid, form, h), 'tlm_create')
  call tlm_check(h, tlm_set_output_capacity(h, int(n_timesteps, c_int32_t)), 'set_output_capacity')
  call tlm_check(h, tlm_set_volume_materials(h, 0_c_int32_t, vm), 'set_volume_materials')
  call tlm_check(h, tlm_set_surface_materials(h, 0_c_int32_t, dummy), 'set_surface_materials')
  call tlm_check(h, tlm_set_cell_codes(h, cell_code, int(n_special, c_int32_t), to_material, to_pml, to_exc, &
                                       to_out, to_cable), 'set_cell_codes')
  call tlm_check(h, tlm_set_face_codes(h, face_code, 0_c_int32_t, dummy, dummy, dummy, dummy), 'set_face_codes')
  call tlm_check(h, tlm_set_excitation_functions(h, 1_c_int32_t, fvalue, fvalue_face), 'set_excitation_functions')
  call tlm_check(h, tlm_set_point_sources(h, 1_c_int32_t, src, 1_c_int32_t, 0_c_int32_t), 'set_point_sources')
  call tlm_check(h, tlm_set_output_points(h, 1_c_int32_t, row, comp), 'set_output_points')

  ! the reference loop body, run_TLM.F90:240-283
  do timestep = 1, n_timesteps
    call gpu_excitation(h, timestep)
    call gpu_scatter(h, timestep)
    call gpu_connect(h, timestep)
    call gpu_outer_boundary(h)
  end do

  allocate (rows(6 * n_timesteps))
  call tlm_check(h, tlm_read_cell_outputs(h, rows, int(n_timesteps, c_int32_t), n_read, first_step), &
                 'read_cell_outputs')
  open (unit=10, file=trim(outname))
  write (10, '(A)') '# Time Domain Data'
  write (10, '(A25,I10)') '# Number of data points: ', 1
  write (10, '(A25,I10)') '# Number of Timesteps  : ', n_timesteps
  do timestep = 1, n_read
    time = (first_step + timestep - 2) * dt
    value = rows(6 * (timestep - 1) + 1)
    if (abs(value) .lt. 1d-30) value = 0d0
    write (10, '(E16.7,I5,E16.7)') time, 1, value
  end do
  close (10)
  call tlm_destroy(h)

end program tlm_gpu_cavity
