/*
 * tlmscn.h -- C ABI of the MI355X (gfx950) SCN scatter/connect engine.
 *
 * One handle holds one TLM solve resident in HBM.  The entry points replace
 * the argument-less Fortran hot-path subroutines of GGI_TLM v2.1.7 and take
 * the reference's own module tables (copied in; the caller keeps ownership).
 * Every table uses the reference's conventions: 1-based row numbers, scan
 * order with x fastest, Fortran column-major for 2-D tables.
 *
 * Reference call sites replaced (SRC/TLM_SOLVER/run_TLM.F90):
 *   CALL excitation()      :242  -> tlm_excitation
 *   CALL scatter()         :244  -> tlm_scatter
 *   CALL cell_output()     :246  -> tlm_read_cell_outputs (batched read)
 *   CALL connect()         :277  -> tlm_connect
 *   CALL outer_boundary()  :283  -> tlm_outer_boundary
 *   the whole loop body    :191-295 -> tlm_step (fused fast path)
 *
 * Error convention: every function returns 0 (TLM_OK) on success and a
 * nonzero TLM_ERR_* code otherwise; tlm_last_error() returns the message.
 * The reference's convention (write(*,*) msg; STOP 1) is applied by the
 * Fortran wrapper module on a nonzero status.
 *
 * Threading: one handle per solve, driven from one host thread; each handle
 * owns its HIP streams (see tlm_set_streams).  Calls on one handle are not
 * re-entrant.
 */
#ifndef TLMSCN_H
#define TLMSCN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TLM_OK 0
#define TLM_ERR_ARG 1         /* bad argument or inconsistent table */
#define TLM_ERR_STATE 2       /* call out of order (e.g. step before tables) */
#define TLM_ERR_HIP 3         /* HIP runtime failure */
#define TLM_ERR_UNSUPPORTED 4 /* feature that stays on the host path */
#define TLM_ERR_NOMEM 5       /* device allocation failed */
#define TLM_ERR_BUDGET 6      /* batch: the solve does not fit the HBM budget left */

/* Numerical forms.  PLAIN: FP64 link planes, bit-faithful to the reference
 * sequential build's operation order.  PRECISION: FP32 link planes; filter
 * states, PML accumulators and the probe DFT stay FP64.  EXTENDED: the
 * precision form's arithmetic and states with each link stored as FP32 plus
 * an 8-bit residual plane (a 32-bit significand, 5 bytes per link): link
 * rounding 256 times smaller than PRECISION's for 25% more link bytes.
 * The batch (tlmbatch.h) takes every form. */
#define TLM_FORM_PLAIN_FP64 0
#define TLM_FORM_PRECISION_FP32 1
#define TLM_FORM_EXTENDED_FP32 2

/* Largest Zfilter order accepted (a and b polynomials of order <= this). */
#define TLM_MAX_FILTER_ORDER 8

/* Volume material types (SRC/TLM_MODULES/TLM_material_modules.F90:72-76). */
#define TLM_VMAT_PEC 1
#define TLM_VMAT_PMC 2
#define TLM_VMAT_DISPERSIVE 3

/* Surface material types (TLM_material_modules.F90:153-169).  PEC, PMC and
 * FREE_SPACE faces are connected on the device; thin layers (DISPERSIVE,
 * ANISOTROPIC_DISPERSIVE), diodes, spice ports and switches are host-path
 * faces whose face voltages the host computes (see tlm_host_faces). */
#define TLM_SMAT_PEC 1
#define TLM_SMAT_PMC 2
#define TLM_SMAT_DISPERSIVE 3
#define TLM_SMAT_FREE_SPACE 4
#define TLM_SMAT_ANISOTROPIC_DISPERSIVE 5
#define TLM_SMAT_DIODE 6
#define TLM_SMAT_SPICE 7
#define TLM_SMAT_SWITCH 9

typedef struct tlm_handle tlm_handle;

/* TLM_general::{nx,ny,nz,n_timesteps,dl,dt,reduced_c_factor,R_*}
 * (SRC/TLM_MODULES/TLM_Modules.F90); one z-slab covering 1..nz. */
typedef struct tlm_grid {
  int32_t nx, ny, nz;
  int32_t n_timesteps; /* length of the excitation function tables */
  double dl, dt;
  double reduced_c_factor;
  double r_xmin, r_xmax, r_ymin, r_ymax, r_zmin, r_zmax;
} tlm_grid;

/* One row of TLM_volume_materials::volume_material_list with the fields the
 * scatter reads (scatter.F90:286-405).  The Zfilters are the slow parts
 * after Z_fast_slow_decomposition (filter_subroutines.F90:1145-1199): a and
 * b coefficients of z^-1, indices 0..order, both of the same order. */
typedef struct tlm_volume_material {
  int32_t type;
  int32_t eps_filter_exists;
  int32_t mu_filter_exists;
  int32_t eps_order;
  int32_t mu_order;
  int32_t pad_;
  double Ys_eps_f, Ge, Zs_mu_f, Rm;
  double eps_a[TLM_MAX_FILTER_ORDER + 1];
  double eps_b[TLM_MAX_FILTER_ORDER + 1];
  double mu_a[TLM_MAX_FILTER_ORDER + 1];
  double mu_b[TLM_MAX_FILTER_ORDER + 1];
} tlm_volume_material;

/* PML_module::PML_material_type without the d_x,d_y,d_z depth fields
 * (SRC/TLM_MODULES/PML_module.F90:102-114). */
typedef struct tlm_pml_parameter {
  double sx, sy, sz;
  double ax, ay, az;
  double exp_x, exp_y, exp_z;
  double Csx, Csy, Csz;
  double Csx2, Csy2, Csz2;
  double Csy_sz, Csz_sx, Csx_sy;
} tlm_pml_parameter;

/* One entry of TLM_excitation::excitation_points, reduced to what
 * Point_excitation (SRC/EXCITATION/excitation_points.F90:231-335) uses.
 * at_face=0: cell-centre source, row = cell_excitation_field_number.
 * at_face=1: face source, row = face_excitation_field_number. */
typedef struct tlm_point_source {
  int32_t function;    /* excitation function number, 1-based */
  int32_t component;   /* 1..12: Ex..Hz, Jx..Jz, Mx..Mz */
  int32_t source_type; /* 0 hard, 1 soft */
  int32_t at_face;
  int32_t row;
} tlm_point_source;

/* One face of TLM_excitation::huygens_surface
 * (SRC/EXCITATION/huygens_surface.F90:734-848). */
typedef struct tlm_huygens_face {
  int32_t row;        /* face_excitation_field_number_list entry; 0 = none */
  int32_t nx, ny, nz; /* outward normal, one of them +-1 */
  double offset;      /* k . r of the face centre */
} tlm_huygens_face;

/* ---- lifecycle ---------------------------------------------------------- */
int tlm_create(const tlm_grid* grid, int32_t form, tlm_handle** out);
void tlm_destroy(tlm_handle* h);
const char* tlm_last_error(const tlm_handle* h);
/* Device bytes held by the handle (link planes, states, tables). */
int64_t tlm_device_bytes(const tlm_handle* h);

/* ---- tables (set once, after tlm_create, before the first step) -------- */

/* mesh::cell_centre_update_code(1:nx*ny*nz) and the cell_update_code_to_*
 * lookup tables (SRC/TLM_MODULES/Mesh_Modules.F90:139-150, built by
 * set_cell_update_codes, SRC/TLM_SOLVER/set_update_codes.F90:38-363, and
 * allocate_volume_material_filter_data, SRC/MATERIALS/
 * TLM_volume_materials.F90:401-518).  to_material is the Fortran (n,2)
 * array, column-major.  A nonzero to_cable entry (the cable junction number,
 * cell_update_code_to_cable_cell_number) makes the cell a host-path cell:
 * see tlm_gather_host_cells.  Host-path cells carry no PML, excitation or
 * output row (TLM_ERR_UNSUPPORTED). */
int tlm_set_cell_codes(tlm_handle* h, const int32_t* cell_centre_update_code,
                       int32_t n_special_cells, const int32_t* to_material,
                       const int32_t* to_pml, const int32_t* to_excitation,
                       const int32_t* to_output, const int32_t* to_cable);

/* mesh::face_update_code(1:number_of_face_codes) and face_update_code_to_*
 * (set_update_codes.F90:383-791).  to_material is the Fortran (n,3) array,
 * column-major; column 1 is the signed surface material number.  A face
 * with a nonzero to_cable entry is connected as its material and excitation
 * make it; its cable junction (connect_normal_to_x.F90:317-319,
 * face_cable_junction) reads and writes only the cable model, and the host
 * runs it after tlm_connect, in the connect's face order. */
int tlm_set_face_codes(tlm_handle* h, const int32_t* face_update_code,
                       int32_t n_special_faces, const int32_t* to_material,
                       const int32_t* to_excitation, const int32_t* to_output,
                       const int32_t* to_cable);

/* TLM_volume_materials::volume_material_list(1:n). */
int tlm_set_volume_materials(tlm_handle* h, int32_t n,
                             const tlm_volume_material* materials);

/* TLM_surface_materials::surface_material_list(1:n)%type. */
int tlm_set_surface_materials(tlm_handle* h, int32_t n, const int32_t* type);

/* PML_module::PML_parameters(1:n_params) and
 * PML_cell_data(1:n_cells)%PML_parameter_array_pos
 * (SRC/MATERIALS/TLM_PML_volume_materials.F90:305-357). */
int tlm_set_pml(tlm_handle* h, int32_t n_params,
                const tlm_pml_parameter* params, int32_t n_cells,
                const int32_t* cell_param);

/* excitation_functions(f)%value(1:n_timesteps) and %value_face, stored as
 * Fortran (n_timesteps, n_functions) arrays, column-major
 * (SRC/EXCITATION/excitation.F90:155-559). */
int tlm_set_excitation_functions(tlm_handle* h, int32_t n_functions,
                                 const double* value,
                                 const double* value_face);

/* Point sources in the reference's excitation_points order.
 * n_cell_rows / n_face_rows = total_number_excitation_cells / _faces. */
int tlm_set_point_sources(tlm_handle* h, int32_t n,
                          const tlm_point_source* sources,
                          int32_t n_cell_rows, int32_t n_face_rows);

/* TLM_excitation::huygens_surface: function number, offset_min, Ei, Hi and
 * the face list (set_huygens_surface_data, huygens_surface.F90:247-715). */
int tlm_set_huygens_surface(tlm_handle* h, int32_t function,
                            double offset_min, const double* Ei,
                            const double* Hi, int32_t n_faces,
                            const tlm_huygens_face* faces);

/* Cell-centre output points: row = output_points(i)%cell_output_field_number,
 * component = output_points(i)%field_component (1..6)
 * (SRC/OUTPUT/output_points.F90:143-237). */
int tlm_set_output_points(tlm_handle* h, int32_t n, const int32_t* row,
                          const int32_t* component);

/* On-the-fly probe DFT frequencies (Hz).  Accumulates, for each output
 * point p and frequency f, sum_n v_p(n) * exp(-j 2 pi f t_n) * dt with
 * t_n = (n-1) dt, the cell-output time of step n, in FP64. */
int tlm_set_dft_frequencies(tlm_handle* h, int32_t n, const double* freq);

/* Frequency-domain outputs (SRC/OUTPUT/frequency_output_volumes.F90:309-386
 * and frequency_output_surfaces.F90:297-421), accumulated on the device:
 * every step in an output's window (first_timestep <= t <= last_timestep,
 * (t - first_timestep) % timestep_interval == 0) adds value * exp(-j 2 pi f
 * time) to each item, time = (t-1) dt for volumes (cell outputs) and
 * (t-1) dt + dt/2 for surfaces (face outputs of the step's connect).
 * kind 1: a volume; its items are cell output rows (1-based, as
 * tlm_set_output_points), value = that row's field_component (1..6).
 * kind 2: a surface; its items are face output rows (1-based), the side of
 * the face whose fields they take (1: the cell that owns the face as its
 * min face, 2: the other) and the outward face normal as a signed axis
 * (+-1, +-2, +-3); field_component 1..6 takes that field, 7..9 the surface
 * current J = n x H (Jx, Jy, Jz).  Items of all outputs are concatenated in
 * rows/sides/normals (sides and normals are read for surfaces only). */
typedef struct {
  double frequency;
  int32_t kind;
  int32_t field_component;
  int32_t first_timestep, last_timestep, timestep_interval;
  int32_t n_items;
} tlm_frequency_output;
int tlm_set_frequency_outputs(tlm_handle* h, int32_t n, const tlm_frequency_output* outputs,
                              const int32_t* rows, const int32_t* sides, const int32_t* normals);

/* Time-domain surface and volume field outputs (SRC/OUTPUT/output_surfaces.F90:
 * 601-948, output_volumes.F90:496-761).  Every step in an output's window
 * (get_output_flag) records one value per item, on the device, as a frame:
 * kind 1: a volume; items are cell output rows, field_component 1..6.
 * kind 2: a surface; items are face output rows with the side of the face
 * whose fields they take and the face normal (signed axis), as for
 * frequency output surfaces; field_component 1..6 a field, 7..9 the
 * surface current J = n x H, 13 |E|, 14 |H|, 15 |J|.  both_sides 1 (the
 * packet's side 0; J components and |J| only) takes J as the jump
 * n x H(side 1) - n x H(side 2) of the face and the fields of side 1.
 * Frames wait on the device, each output in a ring of the handle's output
 * capacity, and move to the host when a ring fills or on read. */
typedef struct {
  int32_t kind;
  int32_t field_component;
  int32_t both_sides;
  int32_t first_timestep, last_timestep, timestep_interval;
  int32_t n_items;
} tlm_field_output;
int tlm_set_field_outputs(tlm_handle* h, int32_t n, const tlm_field_output* outputs, const int32_t* rows,
                          const int32_t* sides, const int32_t* normals);
/* The unread frames of field output k (0-based), oldest first, up to
 * max_frames: values[f * n_items + i].  *n_frames frames are returned and
 * consumed; *first_timestep is the timestep of the first (the others follow
 * at the output's timestep interval). */
int tlm_read_field_output(tlm_handle* h, int32_t k, double* values, int32_t max_frames, int32_t* n_frames,
                          int32_t* first_timestep);

/* Far-field surfaces (SRC/OUTPUT/far_field_surfaces.F90).  The connect of
 * every step adds to each face its surface currents J = n x H and
 * M = -n x E times exp(-j 2 pi f time) dt, time = (t-1) dt + dt/2
 * (face_output_far_field, :252-351), on the device.  tlm_read_far_field
 * transforms them, on the device, to the far field at r = 1 m
 * (write_far_field_surfaces, :370-712): E_theta and E_phi for the angles
 * theta_min + i theta_step (i < n_theta) by phi_min + j phi_step (j < n_phi),
 * theta-major.  dim 3: the 3D transform (face area dl^2); dim 2: the 2D
 * transform in the plane theta = pi/2 (n_theta 1, face length dl).  Faces
 * are given as for frequency output surfaces (face output rows, sides,
 * signed outward normals; n_faces per surface, concatenated), with their
 * centres: 3 coordinates per face in metres (get_cell_point_coordinate). */
typedef struct {
  double frequency;
  double theta_min, theta_step, phi_min, phi_step; /* radians */
  int32_t n_theta, n_phi;
  int32_t dim;
  int32_t n_faces;
} tlm_far_field_surface;
int tlm_set_far_field_surfaces(tlm_handle* h, int32_t n, const tlm_far_field_surface* surfaces,
                               const int32_t* rows, const int32_t* sides, const int32_t* normals,
                               const double* centres);
/* e[4 a + {0, 1, 2, 3}] = Re E_theta, Im E_theta, Re E_phi, Im E_phi at angle
 * a of far-field surface k (0-based). */
int tlm_read_far_field(tlm_handle* h, int32_t k, double* e);

/* RCS surface (SRC/OUTPUT/RCS_surfaces.F90:386-559): the connect of every
 * step adds, per face, the time-domain far field of its currents in the
 * monostatic direction (theta, phi) to far-field point
 *   p = INT(points_per_cell (c0 time - K.r) / dl) - point_offset,
 * K the unit vector of (theta, phi), r the face centre, time = (t-1) dt +
 * dt/2, when 1 <= p <= n_points: E_theta, E_phi, H_theta, H_phi.  The caller
 * sizes n_points and point_offset as initialise_rcs_surfaces does.  Faces
 * as for far-field surfaces.  NULL clears the surface. */
typedef struct {
  double theta, phi; /* radians */
  int32_t n_faces;
  int32_t n_points;
  int32_t point_offset;
  int32_t points_per_cell;
} tlm_rcs_surface;
int tlm_set_rcs_surface(tlm_handle* h, const tlm_rcs_surface* surface, const int32_t* rows,
                        const int32_t* sides, const int32_t* normals, const double* centres);
/* out[4 (p-1) + {0, 1, 2, 3}] = E_theta, E_phi, H_theta, H_phi of point p. */
int tlm_read_rcs(tlm_handle* h, double* out);

/* ---- link state ---------------------------------------------------------- */

/* mesh::V(1:12, 0:nx+1, 0:ny+1, 1:nz), FP64, Fortran layout.  Only the
 * interior cells are read or written; the x/y halo is left untouched. */
int tlm_upload_links(tlm_handle* h, const double* V);
int tlm_download_links(tlm_handle* h, double* V);

/* tlm_download_links without any device allocation.  It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The array is staged through a fixed device buffer of
 * TLM_NOALLOC_STAGING_DOUBLES doubles (32 MiB) one z-slab at a time; a
 * z-plane larger than that (12*(nx+2)*(ny+2) doubles) is refused with
 * TLM_ERR_ARG.  The handle must be set up already (tables uploaded by a
 * step, an upload or tlm_host_cells).  Every call on a GPU shares that GPU's
 * staging buffer: calls that overlap on one GPU race. */
#define TLM_NOALLOC_STAGING_DOUBLES (4 * 1024 * 1024)
int tlm_download_links_noalloc(tlm_handle* h, double* V);

/* ---- stepping (reference call order) ------------------------------------- */
int tlm_excitation(tlm_handle* h, int32_t timestep);
int tlm_scatter(tlm_handle* h, int32_t timestep);
int tlm_connect(tlm_handle* h, int32_t timestep);
int tlm_outer_boundary(tlm_handle* h);

/* Fused fast path: n_steps complete loop bodies starting at first_timestep,
 * equivalent to excitation, scatter, cell_output, connect, outer_boundary
 * for each step. */
int tlm_step(tlm_handle* h, int32_t first_timestep, int32_t n_steps);

/* Wait for all queued work of the handle. */
int tlm_synchronize(tlm_handle* h);

/* ---- outputs ------------------------------------------------------------- */

/* Drain recorded cell_output_field rows: out receives up to max_steps
 * blocks of n_output_cells*6 doubles, row-major (step, cell row, component);
 * *n_read gets the number of steps written, *first_step their first
 * timestep.  Recording is bounded by the ring capacity set here
 * (tlm_set_output_capacity, default 4096 steps). */
int tlm_set_output_capacity(tlm_handle* h, int32_t steps);
int tlm_read_cell_outputs(tlm_handle* h, double* out, int32_t max_steps,
                          int32_t* n_read, int32_t* first_step);

/* Face output rows, face_output_field(row, side, 1:6)
 * (connect_normal_to_x.F90:323-361 and the y, z includes; written each step
 * by face_output_points, output.F90:232-272): one block of
 * n_face_out_rows * 2 sides * 6 fields per connect, rows in face scan order,
 * side 1 = the cell on the face's min side.  Reads the oldest max_steps
 * blocks; *first_step = the timestep of the first connect returned.  A
 * pending connect is applied first, so its block is included. */
int tlm_read_face_outputs(tlm_handle* h, double* out, int32_t max_steps, int32_t* n_read, int32_t* first_step);

/* DFT accumulators: re_im[(p*n_freq + k)*2 + {0,1}]. */
int tlm_read_dft(tlm_handle* h, double* re_im);

/* Accumulators of frequency output k (0-based): re_im[2*i + {0,1}] for its
 * items i in order. */
int tlm_read_frequency_output(tlm_handle* h, int32_t k, double* re_im);

/* ---- host-path cells ------------------------------------------------------
 * Cable junction cells (scatter.F90:452-462, cell_centre_cable_junction)
 * stay on the host: the device pulls their incident voltages (connect and
 * outer boundary) but does not scatter them.  Between tlm_scatter and
 * tlm_connect of a step the host reads their 12 incident voltages with
 * tlm_gather_host_cells, runs its own scatter for those cells (the
 * reference's Thevenin path with the cable junction), and writes the 12
 * scattered voltages back with tlm_scatter_host_cells.  Rows follow the
 * reference's scan order (x fastest); V[12*k + p] is port p+1 of host cell
 * k.  TLM_ERR_STATE outside that window; tlm_step refuses handles with
 * host-path cells (use the split calls). */
int tlm_host_cells(tlm_handle* h, int32_t* n_host, int32_t* cells /* 3*n_host, 1-based, or NULL */);
int tlm_gather_host_cells(tlm_handle* h, double* V);
int tlm_scatter_host_cells(tlm_handle* h, const double* V);

/* ---- host-path faces ------------------------------------------------------
 * Faces of the host-path surface material types (thin layer, anisotropic
 * thin layer, diode, spice port, switch: connect_normal_to_x.F90:75-283 and
 * its y, z counterparts) keep their models on the host.  For such a face
 * the host computes the four face voltages the reference's material branch
 * produces from the four scattered pulses beside the face; the device then
 * applies the face's excitation, completes the connection and records the
 * face output exactly as for any special face (:286-361).  Between
 * tlm_scatter and tlm_connect of a step:
 *   tlm_gather_host_faces  V[4k+0..3] = scattered (a on the face's + side
 *                          cell, a on the - side cell, b on the + side,
 *                          b on the - side), where for a face normal to x
 *                          (a, b) = (Vy, Vz): V(Vy_xmin,cx), V(Vy_xmax,cx-1),
 *                          V(Vz_xmin,cx), V(Vz_xmax,cx-1); normal to y
 *                          (Vx, Vz); normal to z (Vx, Vy);
 *   tlm_scatter_host_faces T[4k+0..3] = (Va_min, Va_max, Vb_min, Vb_max),
 *                          the reference's Vy_min, Vy_max, Vz_min, Vz_max
 *                          (x-normal faces) and likewise for y and z.
 * Rows follow the connect's scan order (cell x fastest; per cell its x-,
 * y-, z-normal min face); tlm_host_faces returns (cx, cy, cz, axis 1..3) of
 * the cell on the face's + side.  tlm_connect refuses a step whose host
 * faces were not written; tlm_step refuses handles with host-path faces. */
int tlm_host_faces(tlm_handle* h, int32_t* n_host_faces, int32_t* faces /* 4*n, 1-based, or NULL */);
int tlm_gather_host_faces(tlm_handle* h, double* V);
int tlm_scatter_host_faces(tlm_handle* h, const double* T);

/* ---- schedule and measurement (no counterpart in the reference) ---------- */

/* Autotuner of the step schedule (no counterpart in the reference): runs
 * steps first_timestep .. first_timestep + *steps_done - 1 of the solve,
 * each candidate schedule (line kernel with and without the XCD mapping,
 * each mixed-line kernel, one or two streams; with TLM_LAYOUT_ROWS also
 * tile kernels of 1/2/4 rows, vector kernel, z-marching kernel: 8
 * candidates, or 14) for steps_per_candidate timed steps after one untimed
 * step, and keeps the fastest.  Every schedule gives the same results, so
 * these are ordinary steps of the solve. */
int tlm_autotune(tlm_handle* h, int32_t first_timestep, int32_t steps_per_candidate, int32_t* steps_done);

/* TLM_set_small_to_zero (TLM_SOLVER/TLM_set_small_to_zero.F90:34-68, gated
 * at run_TLM.F90:236-238; packet read_small_to_zero.F90:55-59): at every
 * step whose number is a multiple of n_small_to_zero, link voltages with
 * abs(V) < small_to_zero_value are set to zero before that step's scatter.
 * n_small_to_zero = 0 turns it off (the default). */
int tlm_set_small_to_zero(tlm_handle* h, int32_t n_small_to_zero, double small_to_zero_value);

/* Storage order of the cells in the link planes (no counterpart in the
 * reference, whose V array is scan-ordered; results do not depend on it).
 * TLM_LAYOUT_X_SEGMENTS (default): the x range is cut where at least an
 * eighth of an x-plane's cells change class between neighbouring planes
 * (PML slabs, Huygens planes, walls), at most TLM_MAX_SEGMENTS segments;
 * segment s stores its cells x-fastest with its own row width, so a PML
 * slab or a source plane fills whole 128-byte lines by itself.
 * TLM_LAYOUT_ROWS: one segment, the reference's scan order.  Set before the
 * first call that builds the device tables (step, scatter, upload, kernel
 * byte query); TLM_ERR_STATE after. */
#define TLM_LAYOUT_X_SEGMENTS 0
#define TLM_LAYOUT_ROWS 1
#define TLM_MAX_SEGMENTS 16
int tlm_set_layout(tlm_handle* h, int32_t layout);

/* The layout in use (builds the device tables): *n_segments segments,
 * segment s holding x in [x0[s], x0[s+1]) (0-based); x0 has room for
 * TLM_MAX_SEGMENTS + 1 entries, or is NULL. */
int tlm_layout_segments(tlm_handle* h, int32_t* n_segments, int32_t* x0);

/* Bulk-kernel schedule: variant 3 (default) = 256 consecutive storage cells
 * per workgroup over the pure bulk lines; variants 0-2 need TLM_LAYOUT_ROWS:
 * 0 = one cell per lane, 64 x rows tiles; 1 = workgroups marching zc planes;
 * 2 = 16-byte vector accesses, several consecutive cells per lane (falls
 * back to 0 when nx is not a multiple of the vector width); rows = tile
 * height (1, 2 or 4 rows); xcd_remap = 1: give each XCD a contiguous range
 * of workgroups; 2 (line kernel only): XCD k takes the blocks of the k-th
 * y-band of every z-plane in storage order, so the y-1 and z-1 neighbour
 * rows of a block were read by the same XCD just before (its own L2).
 * mixed: mixed-line kernel, 1 = LDS-staged cell-order loads and stores
 * around class-sorted node updates, 0 = class-sorted lanes with four bulk
 * cells per lane, -1 = the default (LDS-staged).
 * Any schedule gives identical results; it only changes speed
 * (tlm_autotune picks one by timing). */
int tlm_set_schedule(tlm_handle* h, int32_t variant, int32_t rows, int32_t zc, int32_t xcd_remap,
                     int32_t mixed);

/* The schedule in use (after tlm_autotune: the one it kept); any output
 * pointer may be NULL. */
int tlm_get_schedule(const tlm_handle* h, int32_t* variant, int32_t* rows, int32_t* zc, int32_t* xcd_remap,
                     int32_t* mixed, int32_t* streams);

/* Streams of a step (no counterpart in the reference): 1 = every kernel of
 * a step on the handle's stream in turn; 2 (default) = the material,
 * generic and PML-material class kernels on a second stream, concurrent
 * with the bulk, PML and mixed-line kernels (they write disjoint cells and
 * read only the previous step's buffer), joined before the step ends.
 * Profiling serialises them.  Results do not depend on it. */
int tlm_set_streams(tlm_handle* h, int32_t streams);

/* Kernel classes timed by the profiling hooks. */
#define TLM_KCLASS_BULK 0
#define TLM_KCLASS_PML 1
#define TLM_KCLASS_GENERIC 2
#define TLM_KCLASS_DFT 3
#define TLM_KCLASS_MATERIAL 4
#define TLM_KCLASS_MIXED 5
#define TLM_KCLASS_PML_MATERIAL 6
#define TLM_N_KCLASS 7

/* When enabled, every kernel launch of a step is bracketed by HIP events
 * recorded on the handle's own stream.  tlm_read_kernel_times waits for the
 * stream and returns, per class, the launch count and the summed duration
 * (ms) since the previous read, then resets them. */
int tlm_set_profiling(tlm_handle* h, int32_t enable);
int tlm_read_kernel_times(tlm_handle* h, int64_t* launches, double* ms);

/* Per-class cell counts and algorithmic HBM bytes per launch (links read
 * and written, FP64 state words read and written, per-cell indices), the
 * figures the roofline fraction is computed from. */
int tlm_kernel_bytes(tlm_handle* h, int64_t* cells, int64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* TLMSCN_H */
