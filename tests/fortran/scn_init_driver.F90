!
! Test driver of gpu_scn_init: fills the reference's module globals (the
! state GGI_TLM's set-up leaves in TLM_general, mesh, PML_module,
! TLM_volume_materials, TLM_surface_materials, TLM_excitation, TLM_output)
! from a table file written by tests/test_gpu_scn_init.py, calls
! gpu_scn_init, steps the solve and writes the cell-output rows.
!   scn_init_driver <tables file> <output file> <steps>
!
program scn_init_driver

  use iso_c_binding
  use tlm_gpu_iface
  use gpu_scn
  use TLM_general
  use mesh
  use PML_module
  use TLM_volume_materials
  use TLM_surface_materials
  use TLM_excitation
  use TLM_output

  implicit none

  character(len=1024) :: fin, fout, arg
  integer :: u, i, k, n, m, steps, nf, nrows, o
  integer :: ival(6)
  real*8 :: rval(6)
  type(c_ptr) :: h
  integer(c_int32_t) :: n_read, first
  real(c_double), allocatable :: out(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg)
  read(arg, *) steps
  open(newunit=u, file=trim(fin), status='old', action='read')

  ! TLM_general
  read(u, *) nx, ny, nz, n_timesteps
  read(u, *) dl, dt, reduced_c_factor
  read(u, *) R_xmin, R_xmax, R_ymin, R_ymax, R_zmin, R_zmax
  nz1 = 1
  nz2 = nz
  set_small_to_zero = .false.

  ! mesh: cell codes and lookup tables
  number_of_cell_centre_codes = nx * ny * nz
  allocate(cell_centre_update_code(number_of_cell_centre_codes))
  read(u, *) cell_centre_update_code
  read(u, *) n_special_cells
  n = max(n_special_cells, 1)
  allocate(cell_update_code_to_material_data(n, 2), cell_update_code_to_PML_data(n),             &
           cell_update_code_to_excitation_number(n), cell_update_code_to_output_number(n),       &
           cell_update_code_to_cable_cell_number(n))
  cell_update_code_to_material_data = 0
  cell_update_code_to_PML_data = 0
  cell_update_code_to_excitation_number = 0
  cell_update_code_to_output_number = 0
  cell_update_code_to_cable_cell_number = 0
  if (n_special_cells .gt. 0) then
    read(u, *) cell_update_code_to_material_data(1:n_special_cells, 1)
    read(u, *) cell_update_code_to_material_data(1:n_special_cells, 2)
    read(u, *) cell_update_code_to_PML_data(1:n_special_cells)
    read(u, *) cell_update_code_to_excitation_number(1:n_special_cells)
    read(u, *) cell_update_code_to_output_number(1:n_special_cells)
    read(u, *) cell_update_code_to_cable_cell_number(1:n_special_cells)
  end if

  ! mesh: face codes and lookup tables
  read(u, *) number_of_face_codes
  allocate(face_update_code(max(number_of_face_codes, 1)))
  face_update_code = 0
  if (number_of_face_codes .gt. 0) read(u, *) face_update_code(1:number_of_face_codes)
  read(u, *) n_special_faces
  n = max(n_special_faces, 1)
  allocate(face_update_code_to_material_data(n, 3), face_update_code_to_excitation_number(n),     &
           face_update_code_to_output_number(n), face_update_code_to_cable_number(n))
  face_update_code_to_material_data = 0
  face_update_code_to_excitation_number = 0
  face_update_code_to_output_number = 0
  face_update_code_to_cable_number = 0
  if (n_special_faces .gt. 0) then
    do k = 1, 3
      read(u, *) face_update_code_to_material_data(1:n_special_faces, k)
    end do
    read(u, *) face_update_code_to_excitation_number(1:n_special_faces)
    read(u, *) face_update_code_to_output_number(1:n_special_faces)
    read(u, *) face_update_code_to_cable_number(1:n_special_faces)
  end if

  ! TLM_volume_materials
  read(u, *) n_volume_materials
  allocate(volume_material_list(max(n_volume_materials, 1)))
  do m = 1, n_volume_materials
    read(u, *) ival(1:5)
    read(u, *) rval(1:4)
    volume_material_list(m)%type = ival(1)
    volume_material_list(m)%eps_filter_exists = ival(2) .ne. 0
    volume_material_list(m)%mu_filter_exists = ival(3) .ne. 0
    volume_material_list(m)%Ys_eps_f = rval(1)
    volume_material_list(m)%Ge = rval(2)
    volume_material_list(m)%Zs_mu_f = rval(3)
    volume_material_list(m)%Rm = rval(4)
    o = ival(4)
    volume_material_list(m)%Zs_eps_Z%a%order = o
    volume_material_list(m)%Zs_eps_Z%b%order = o
    allocate(volume_material_list(m)%Zs_eps_Z%a%coeff(0:o), volume_material_list(m)%Zs_eps_Z%b%coeff(0:o))
    read(u, *) volume_material_list(m)%Zs_eps_Z%a%coeff(0:o)
    read(u, *) volume_material_list(m)%Zs_eps_Z%b%coeff(0:o)
    o = ival(5)
    volume_material_list(m)%Zs_mu_Z%a%order = o
    volume_material_list(m)%Zs_mu_Z%b%order = o
    allocate(volume_material_list(m)%Zs_mu_Z%a%coeff(0:o), volume_material_list(m)%Zs_mu_Z%b%coeff(0:o))
    read(u, *) volume_material_list(m)%Zs_mu_Z%a%coeff(0:o)
    read(u, *) volume_material_list(m)%Zs_mu_Z%b%coeff(0:o)
  end do

  ! TLM_surface_materials
  read(u, *) n_surface_materials
  allocate(surface_material_list(max(n_surface_materials, 1)))
  do m = 1, n_surface_materials
    read(u, *) surface_material_list(m)%type
  end do

  ! PML_module
  read(u, *) PML_n_parameters, total_number_of_PML_cells
  PML_flag = total_number_of_PML_cells .gt. 0
  allocate(PML_parameters(max(PML_n_parameters, 1)), PML_cell_data(max(total_number_of_PML_cells, 1)))
  do k = 1, PML_n_parameters
    read(u, *) PML_parameters(k)%sx, PML_parameters(k)%sy, PML_parameters(k)%sz,                  &
               PML_parameters(k)%ax, PML_parameters(k)%ay, PML_parameters(k)%az,                  &
               PML_parameters(k)%exp_x, PML_parameters(k)%exp_y, PML_parameters(k)%exp_z,         &
               PML_parameters(k)%Csx, PML_parameters(k)%Csy, PML_parameters(k)%Csz,               &
               PML_parameters(k)%Csx2, PML_parameters(k)%Csy2, PML_parameters(k)%Csz2,            &
               PML_parameters(k)%Csy_sz, PML_parameters(k)%Csz_sx, PML_parameters(k)%Csx_sy
  end do
  do i = 1, total_number_of_PML_cells
    read(u, *) PML_cell_data(i)%PML_parameter_array_pos
  end do

  ! TLM_excitation
  read(u, *) n_excitation_functions
  allocate(excitation_functions(max(n_excitation_functions, 1)))
  do k = 1, n_excitation_functions
    allocate(excitation_functions(k)%value(n_timesteps), excitation_functions(k)%value_face(n_timesteps))
    read(u, *) excitation_functions(k)%value
    read(u, *) excitation_functions(k)%value_face
  end do
  read(u, *) n_excitation_points, total_number_excitation_cells, total_number_excitation_faces
  allocate(excitation_points(max(n_excitation_points, 1)))
  do k = 1, n_excitation_points
    read(u, *) ival(1:5)
    excitation_points(k)%excitation_function_number = ival(1)
    excitation_points(k)%field_component = ival(2)
    excitation_points(k)%source_type = ival(3)
    excitation_points(k)%cell_excitation_field_number = 0
    excitation_points(k)%face_excitation_field_number = 0
    if (ival(4) .eq. 0) then
      excitation_points(k)%cell_excitation_field_number = ival(5)
    else
      excitation_points(k)%face_excitation_field_number = ival(5)
    end if
  end do
  n_excitation_surfaces = 0
  n_excitation_volumes = 0
  n_excitation_modes = 0
  read(u, *) n
  n_huygens_surfaces = 0
  if (n .gt. 0) then
    n_huygens_surfaces = 1
    huygens_surface%outer_surface_flag = .false.
    huygens_surface%n_surface_patches = n
    read(u, *) huygens_surface%excitation_function_number, huygens_surface%offset_min
    read(u, *) huygens_surface%Ei, huygens_surface%Hi
    allocate(huygens_surface%face_excitation_field_number_list(n), huygens_surface%nx(n),           &
             huygens_surface%ny(n), huygens_surface%nz(n), huygens_surface%offset(n))
    do k = 1, n
      read(u, *) huygens_surface%face_excitation_field_number_list(k), huygens_surface%nx(k),       &
                 huygens_surface%ny(k), huygens_surface%nz(k), huygens_surface%offset(k)
    end do
  end if

  ! TLM_output
  read(u, *) n_output_points, nrows
  allocate(output_points(max(n_output_points, 1)))
  do k = 1, n_output_points
    read(u, *) output_points(k)%cell_output_field_number, output_points(k)%field_component
    output_points(k)%face_output_field_number = 0
  end do
  close(u)

  call gpu_scn_init(h, TLM_FORM_PLAIN_FP64, int(steps, c_int32_t))
  call tlm_check(h, tlm_step(h, 1_c_int32_t, int(steps, c_int32_t)), 'tlm_step')
  allocate(out(6 * max(nrows, 1) * steps))
  out = 0d0
  call tlm_check(h, tlm_read_cell_outputs(h, out, int(steps, c_int32_t), n_read, first), 'read_cell_outputs')
  open(newunit=u, file=trim(fout), status='replace', action='write')
  write(u, *) n_read, first
  do i = 1, 6 * nrows * n_read
    write(u, '(ES25.17E3)') out(i)
  end do
  close(u)
  call tlm_destroy(h)
end program scn_init_driver
