!
! gpu_scn_init.F90 -- hands GGI_TLM's set-up tables to the MI355X engine.
!
! Compiled inside the reference's build (it uses the reference's modules:
! SRC/TLM_MODULES/{TLM_Modules,Mesh_Modules,PML_module,
! TLM_material_modules,TLM_excitation_modules,TLM_output_modules}.F90).
! Called once after the set-up (GGI_TLM.F90:142-152: mesh, update codes,
! materials, PML, excitation, outputs) and before the time loop of
! run_TLM.F90, it flattens the module globals into the engine's bind(c)
! tables (ggi_tlm_amd/fortran/tlm_gpu_iface.F90, include/tlmscn.h):
!
!   TLM_general          -> tlm_grid, small-to-zero packet
!   mesh                 -> cell and face update codes and their lookup
!                           tables (Mesh_Modules.F90:98-150)
!   TLM_volume_materials -> volume_material_list rows with the slow
!                           Zfilter coefficients (TLM_material_modules.F90:42-70,
!                           filter_types.F90:45-62)
!   TLM_surface_materials-> surface material types
!   PML_module           -> PML_parameters and each PML cell's parameter row
!                           (PML_module.F90:102-156)
!   TLM_excitation       -> excitation function tables, point, surface and
!                           volume sources as point-source rows, the Huygens
!                           surface (TLM_excitation_modules.F90:160-171)
!   TLM_output           -> cell-centre output points
!
! Errors follow the reference's convention (write(*,*) message; STOP 1).
!
module gpu_scn

  use iso_c_binding
  use tlm_gpu_iface
  use TLM_general
  use mesh
  use PML_module
  use TLM_volume_materials
  use TLM_surface_materials
  use TLM_excitation
  use TLM_output

  implicit none

contains

  subroutine gpu_scn_check(h, rc, what)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc .ne. 0) then
      write(*,*) 'ERROR in gpu_scn_init: ', what, ' returned ', rc
      if (c_associated(h)) call tlm_check(h, rc, what)
      stop 1
    end if
  end subroutine gpu_scn_check

  ! One Zfilter (a(z)/b(z), polynomials in z^-1) into coefficient rows of
  ! the common order max(a%order, b%order), zero padded.
  subroutine zfilter_row(zf, order, a, b)
    type(Zfilter), intent(in) :: zf
    integer(c_int32_t), intent(out) :: order
    real(c_double), intent(out) :: a(0:TLM_MAX_FILTER_ORDER), b(0:TLM_MAX_FILTER_ORDER)
    integer :: oa, ob
    a = 0d0
    b = 0d0
    oa = zf%a%order
    ob = zf%b%order
    order = max(oa, ob)
    if (order .gt. TLM_MAX_FILTER_ORDER) then
      write(*,*) 'ERROR in gpu_scn_init: filter order', order, ' exceeds', TLM_MAX_FILTER_ORDER
      stop 1
    end if
    a(0:oa) = zf%a%coeff(0:oa)
    b(0:ob) = zf%b%coeff(0:ob)
  end subroutine zfilter_row

  ! Create the engine handle h in numerical form `form` and load every
  ! table.  output_capacity: timesteps of cell-output rows kept on the
  ! device between reads (tlm_set_output_capacity).
  subroutine gpu_scn_init(h, form, output_capacity)
    type(c_ptr), intent(out) :: h
    integer(c_int32_t), intent(in) :: form, output_capacity

    type(tlm_grid) :: g
    type(tlm_volume_material), allocatable :: vm(:)
    type(tlm_pml_parameter), allocatable :: pp(:)
    type(tlm_point_source), allocatable :: ps(:)
    type(tlm_huygens_face), allocatable :: hf(:)
    integer(c_int32_t), allocatable :: ibuf(:), orow(:), ocomp(:)
    integer(c_int32_t), allocatable :: cm(:, :), fm(:, :)
    real(c_double), allocatable :: val(:, :), valf(:, :)
    integer :: i, k, n, m, nf, nsrc

    h = c_null_ptr
    if (nz1 .ne. 1 .or. nz2 .ne. nz) then
      write(*,*) 'ERROR in gpu_scn_init: the engine runs the whole mesh (one z-slab) on one GPU'
      stop 1
    end if

    ! ---- grid (TLM_general) ------------------------------------------------
    g%nx = nx
    g%ny = ny
    g%nz = nz
    g%n_timesteps = n_timesteps
    g%dl = dl
    g%dt = dt
    g%reduced_c_factor = reduced_c_factor
    g%r_xmin = R_xmin
    g%r_xmax = R_xmax
    g%r_ymin = R_ymin
    g%r_ymax = R_ymax
    g%r_zmin = R_zmin
    g%r_zmax = R_zmax
    call gpu_scn_check(h, tlm_create(g, form, h), 'tlm_create')

    ! ---- volume materials (before the cell codes, which refer to them) ----
    n = n_volume_materials
    allocate(vm(max(n, 1)))
    do m = 1, n
      vm(m)%mtype = volume_material_list(m)%type
      vm(m)%eps_filter_exists = 0
      vm(m)%mu_filter_exists = 0
      vm(m)%eps_order = 0
      vm(m)%mu_order = 0
      vm(m)%pad_ = 0
      vm(m)%Ys_eps_f = 0d0
      vm(m)%Ge = 0d0
      vm(m)%Zs_mu_f = 0d0
      vm(m)%Rm = 0d0
      vm(m)%eps_a = 0d0
      vm(m)%eps_b = 0d0
      vm(m)%mu_a = 0d0
      vm(m)%mu_b = 0d0
      vm(m)%eps_b(0) = 1d0
      vm(m)%mu_b(0) = 1d0
      if (volume_material_list(m)%type .eq. volume_material_type_DISPERSIVE) then
        vm(m)%Ge = volume_material_list(m)%Ge
        vm(m)%Rm = volume_material_list(m)%Rm
        if (volume_material_list(m)%eps_filter_exists) then
          vm(m)%eps_filter_exists = 1
          vm(m)%Ys_eps_f = volume_material_list(m)%Ys_eps_f
          call zfilter_row(volume_material_list(m)%Zs_eps_Z, vm(m)%eps_order, vm(m)%eps_a, vm(m)%eps_b)
        end if
        if (volume_material_list(m)%mu_filter_exists) then
          vm(m)%mu_filter_exists = 1
          vm(m)%Zs_mu_f = volume_material_list(m)%Zs_mu_f
          call zfilter_row(volume_material_list(m)%Zs_mu_Z, vm(m)%mu_order, vm(m)%mu_a, vm(m)%mu_b)
        end if
      end if
    end do
    call gpu_scn_check(h, tlm_set_volume_materials(h, int(n, c_int32_t), vm), 'tlm_set_volume_materials')

    ! ---- surface materials --------------------------------------------------
    n = n_surface_materials
    allocate(ibuf(max(n, 1)))
    ibuf = 0
    do m = 1, n
      ibuf(m) = surface_material_list(m)%type
    end do
    call gpu_scn_check(h, tlm_set_surface_materials(h, int(n, c_int32_t), ibuf), 'tlm_set_surface_materials')
    deallocate(ibuf)

    ! ---- cell and face update codes (mesh) ----------------------------------
    n = n_special_cells
    allocate(cm(max(n, 1), 2))
    cm = 0
    if (n .gt. 0) cm(1:n, 1:2) = cell_update_code_to_material_data(1:n, 1:2)
    call gpu_scn_check(h, tlm_set_cell_codes(h, cell_centre_update_code, int(n, c_int32_t), cm,     &
                       cell_update_code_to_PML_data, cell_update_code_to_excitation_number,         &
                       cell_update_code_to_output_number, cell_update_code_to_cable_cell_number),  &
                       'tlm_set_cell_codes')
    n = n_special_faces
    allocate(fm(max(n, 1), 3))
    fm = 0
    if (n .gt. 0) fm(1:n, 1:3) = face_update_code_to_material_data(1:n, 1:3)
    call gpu_scn_check(h, tlm_set_face_codes(h, face_update_code, int(n, c_int32_t), fm,            &
                       face_update_code_to_excitation_number, face_update_code_to_output_number,   &
                       face_update_code_to_cable_number), 'tlm_set_face_codes')

    ! ---- PML (PML_module) ---------------------------------------------------
    if (PML_flag .and. total_number_of_PML_cells .gt. 0) then
      n = PML_n_parameters
      allocate(pp(n))
      do k = 1, n
        pp(k)%sx = PML_parameters(k)%sx
        pp(k)%sy = PML_parameters(k)%sy
        pp(k)%sz = PML_parameters(k)%sz
        pp(k)%ax = PML_parameters(k)%ax
        pp(k)%ay = PML_parameters(k)%ay
        pp(k)%az = PML_parameters(k)%az
        pp(k)%exp_x = PML_parameters(k)%exp_x
        pp(k)%exp_y = PML_parameters(k)%exp_y
        pp(k)%exp_z = PML_parameters(k)%exp_z
        pp(k)%Csx = PML_parameters(k)%Csx
        pp(k)%Csy = PML_parameters(k)%Csy
        pp(k)%Csz = PML_parameters(k)%Csz
        pp(k)%Csx2 = PML_parameters(k)%Csx2
        pp(k)%Csy2 = PML_parameters(k)%Csy2
        pp(k)%Csz2 = PML_parameters(k)%Csz2
        pp(k)%Csy_sz = PML_parameters(k)%Csy_sz
        pp(k)%Csz_sx = PML_parameters(k)%Csz_sx
        pp(k)%Csx_sy = PML_parameters(k)%Csx_sy
      end do
      allocate(ibuf(total_number_of_PML_cells))
      do i = 1, total_number_of_PML_cells
        ibuf(i) = PML_cell_data(i)%PML_parameter_array_pos
      end do
      call gpu_scn_check(h, tlm_set_pml(h, int(n, c_int32_t), pp, int(total_number_of_PML_cells, c_int32_t), &
                         ibuf), 'tlm_set_pml')
      deallocate(ibuf)
    end if

    ! ---- excitation (TLM_excitation) ----------------------------------------
    nf = n_excitation_functions
    allocate(val(max(n_timesteps, 1), max(nf, 1)), valf(max(n_timesteps, 1), max(nf, 1)))
    val = 0d0
    valf = 0d0
    do k = 1, nf
      val(1:n_timesteps, k) = excitation_functions(k)%value(1:n_timesteps)
      valf(1:n_timesteps, k) = excitation_functions(k)%value_face(1:n_timesteps)
    end do
    call gpu_scn_check(h, tlm_set_excitation_functions(h, int(nf, c_int32_t), val, valf), &
                       'tlm_set_excitation_functions')
    ! point, surface and volume sources as point-source rows, in the order
    ! excitation() applies them (points, surfaces, volumes)
    nsrc = n_excitation_points
    do k = 1, n_excitation_surfaces
      nsrc = nsrc + excitation_surfaces(k)%number_of_faces
    end do
    do k = 1, n_excitation_volumes
      nsrc = nsrc + excitation_volumes(k)%number_of_cells
    end do
    if (n_excitation_modes .gt. 0) then
      write(*,*) 'ERROR in gpu_scn_init: mode excitations stay on the host path'
      stop 1
    end if
    allocate(ps(max(nsrc, 1)))
    n = 0
    do k = 1, n_excitation_points
      n = n + 1
      ps(n)%func = excitation_points(k)%excitation_function_number
      ps(n)%component = excitation_points(k)%field_component
      ps(n)%source_type = excitation_points(k)%source_type
      if (excitation_points(k)%cell_excitation_field_number .ne. 0) then
        ps(n)%at_face = 0
        ps(n)%row = excitation_points(k)%cell_excitation_field_number
      else
        ps(n)%at_face = 1
        ps(n)%row = excitation_points(k)%face_excitation_field_number
      end if
    end do
    do k = 1, n_excitation_surfaces
      do i = 1, excitation_surfaces(k)%number_of_faces
        n = n + 1
        ps(n)%func = excitation_surfaces(k)%excitation_function_number
        ps(n)%component = excitation_surfaces(k)%field_component
        ps(n)%source_type = excitation_surfaces(k)%source_type
        ps(n)%at_face = 1
        ps(n)%row = excitation_surfaces(k)%face_excitation_field_number_list(i)
      end do
    end do
    do k = 1, n_excitation_volumes
      do i = 1, excitation_volumes(k)%number_of_cells
        n = n + 1
        ps(n)%func = excitation_volumes(k)%excitation_function_number
        ps(n)%component = excitation_volumes(k)%field_component
        ps(n)%source_type = excitation_volumes(k)%source_type
        ps(n)%at_face = 0
        ps(n)%row = excitation_volumes(k)%cell_excitation_field_number_list(i)
      end do
    end do
    call gpu_scn_check(h, tlm_set_point_sources(h, int(n, c_int32_t), ps, int(total_number_excitation_cells, &
                       c_int32_t), int(total_number_excitation_faces, c_int32_t)), 'tlm_set_point_sources')
    if (n_huygens_surfaces .gt. 0) then
      if (huygens_surface%outer_surface_flag) then
        write(*,*) 'ERROR in gpu_scn_init: a Huygens surface on the outer boundary stays on the host path'
        stop 1
      end if
      n = huygens_surface%n_surface_patches
      allocate(hf(max(n, 1)))
      do k = 1, n
        hf(k)%row = huygens_surface%face_excitation_field_number_list(k)
        hf(k)%nx = huygens_surface%nx(k)
        hf(k)%ny = huygens_surface%ny(k)
        hf(k)%nz = huygens_surface%nz(k)
        hf(k)%offset = huygens_surface%offset(k)
      end do
      call gpu_scn_check(h, tlm_set_huygens_surface(h, int(huygens_surface%excitation_function_number, c_int32_t), &
                         huygens_surface%offset_min, huygens_surface%Ei, huygens_surface%Hi, int(n, c_int32_t), hf), &
                         'tlm_set_huygens_surface')
    end if

    ! ---- outputs (TLM_output): cell-centre output points ---------------------
    n = 0
    do k = 1, n_output_points
      if (output_points(k)%cell_output_field_number .ne. 0) n = n + 1
    end do
    allocate(orow(max(n, 1)), ocomp(max(n, 1)))
    n = 0
    do k = 1, n_output_points
      if (output_points(k)%cell_output_field_number .eq. 0) cycle
      n = n + 1
      orow(n) = output_points(k)%cell_output_field_number
      ocomp(n) = output_points(k)%field_component
    end do
    call gpu_scn_check(h, tlm_set_output_points(h, int(n, c_int32_t), orow, ocomp), 'tlm_set_output_points')
    call gpu_scn_check(h, tlm_set_output_capacity(h, output_capacity), 'tlm_set_output_capacity')

    ! ---- TLM_set_small_to_zero packet ---------------------------------------
    if (set_small_to_zero) then
      call gpu_scn_check(h, tlm_set_small_to_zero(h, int(n_small_to_zero, c_int32_t), small_to_zero_value), &
                         'tlm_set_small_to_zero')
    end if
  end subroutine gpu_scn_init

end module gpu_scn
