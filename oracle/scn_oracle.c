/*
 * scn_oracle.c -- CPU oracle for the SCN scatter/connect hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see scn_oracle.h).  A single-threaded FP64
 * restatement of the GGI_TLM v2.1.7 loop body, kept in the reference's data
 * layout (V(1:12,0:nx+1,0:ny+1,1:nz), column-major, 12 links of a cell
 * contiguous) and scan order (x fastest), with every expression written in
 * the reference's evaluation order.  Build with -ffp-contract=off so no
 * multiply-add is fused.
 *
 * Reference files followed (paths under GGI_TLM/SRC):
 *   TLM_SOLVER/scatter.F90:216-585        scatter, free space and Thevenin
 *   TLM_SOLVER/scatter.F90:606-661        add_shunt_admittance / series
 *   TLM_SOLVER/scatter_PML.F90:22-313     PML cell update (stages 1-13)
 *   TLM_SOLVER/scatter_PML_material.F90:22-534  PML cell holding a material
 *   TLM_SOLVER/TLM_set_small_to_zero.F90:34-68   small-value flush
 *   FILTER/filter_subroutines.F90:510-575 evaluate_Zfilter / timeshift
 *   TLM_SOLVER/connect.F90:180-214        face loop
 *   TLM_SOLVER/connect_normal_to_{x,y,z}.F90  face update, sources, outputs
 *   TLM_SOLVER/outer_boundary.F90:71-131  reflection at the mesh boundary
 *   EXCITATION/excitation.F90:578-626     per-step excitation reset
 *   EXCITATION/excitation_points.F90:231-335  point sources
 *   EXCITATION/huygens_surface.F90:734-848    Huygens face sources
 *   EXCITATION/excitation.F90:645-690     get_interpolated_excitation_value
 *   OUTPUT/output_points.F90:463-612      cell output capture
 */
#include "scn_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* SRC/CONSTANTS/constants.F90:27-39 */
static const double PI = 3.141592653589793;
static const double Z0 = 376.73031346177;
static const double C0 = 2.99792458e8;
#define Y0 (1.0 / Z0)

/* Port numbering, SRC/TLM_MODULES/Mesh_Modules.F90:83-94 (1-based). */
enum {
  Vy_xmin = 1, Vz_xmin, Vy_xmax, Vz_xmax, Vx_ymin, Vz_ymin,
  Vx_ymax, Vz_ymax, Vx_zmin, Vy_zmin, Vx_zmax, Vy_zmax
};
/* Field components, SRC/TLM_MODULES/TLM_Modules.F90:164-176 (1-based). */
enum { Ex = 1, Ey, Ez, Hx, Hy, Hz, Jx, Jy, Jz, Mx, My, Mz };

#define NW (TLM_MAX_FILTER_ORDER + 1)

struct oracle_handle {
  tlm_grid g;
  int nx, ny, nz;
  char err[512];
  double* V; /* 12*(nx+2)*(ny+2)*nz */
  /* cell codes */
  int32_t* cell_code;
  int n_special;
  int32_t *c_mat, *c_filt, *c_pml, *c_exc, *c_out, *c_cable;
  int n_out_rows, n_exc_rows;
  /* face codes */
  int32_t* face_code;
  int n_faces, n_special_faces;
  int32_t *f_mat, *f_exc, *f_out;
  int n_face_out_rows;
  /* materials */
  int n_vmat;
  tlm_volume_material* vmat;
  int n_smat;
  int32_t* smat;
  /* Zfilter states w(0:order), per filter number (1-based) */
  int n_filters;
  double *eps_w, *mu_w;
  int32_t *eps_ord, *mu_ord;
  /* PML */
  int n_pml_params, n_pml_cells;
  tlm_pml_parameter* pml_par;
  int32_t* pml_cell_param;
  double* pml_state; /* per PML cell: 27 words, see PS_* */
  /* PML cells holding a (non-dispersive) material: PML_material_array_pos
   * per PML cell (0 = none) and PML_material_data_type words, see PM_* */
  int n_pml_mat;
  int32_t* pml_mat_pos;
  double* pml_mat_state;
  /* set_small_to_zero packet: n_small_to_zero (0: off), small_to_zero_value */
  int stz_n;
  int layout; /* tlm_set_layout's value: decides the autotuner's candidate count only */
  double stz_value;
  /* excitation */
  int n_functions;
  double *value, *value_face;
  int n_ps;
  tlm_point_source* ps;
  int n_cell_rows, n_face_rows;
  double* cell_exc_field; /* (row, 1:9) row-major [row*9 + c-1] */
  int32_t* cell_exc_type;  /* (row, 1:6) */
  double* face_exc_field;  /* (row, side, 1:12) [(row*2+side)*12 + c-1] */
  int32_t* face_exc_type;  /* (row, side, 1:6) */
  int huy_function;
  double huy_offset_min, Ei[3], Hi[3];
  int n_huy;
  tlm_huygens_face* huy;
  /* outputs */
  double* cell_out; /* (row, 1:6) */
  double* face_out; /* (row, side, 1:6) */
  int fring_count, fring_first;
  double* fring; /* cap blocks of n_face_out_rows*12, one per connect */
  int n_points;
  int32_t *pt_row, *pt_comp;
  int n_freq;
  double* freq;
  double* dft; /* (point, freq, 2) */
  /* frequency-domain outputs: per item its output, row, side, normal and
     complex accumulator */
  int n_fo, n_fo_items;
  tlm_frequency_output* fo;
  int32_t *fo_out, *fo_row, *fo_side, *fo_norm;
  double* fo_acc;
  /* time-domain field outputs: per output its items and the recorded
     frames not yet read (values, count, timestep of the first) */
  int n_fd;
  tlm_field_output* fd;
  int32_t *fd_row, *fd_side, *fd_norm; /* items of all outputs, concatenated */
  double** fd_frames;
  int *fd_nframes, *fd_capframes, *fd_first;
  /* far-field surfaces: per surface its faces (row, side, normal, centre)
     and the accumulated J(face, 1:3), M(face, 1:3) as complex pairs */
  int n_ff;
  tlm_far_field_surface* ff;
  int32_t *ff_row, *ff_side, *ff_norm;
  double *ff_centre, *ff_J, *ff_M;
  /* RCS surface: faces in the caller's order and the far-field series
     E_theta, E_phi, H_theta, H_phi per point */
  int rcs_on;
  tlm_rcs_surface rcs;
  int32_t *rcs_row, *rcs_side, *rcs_norm;
  double *rcs_centre, *rcs_E;
  int cap, ring_count, ring_first;
  double* ring; /* cap blocks of n_out_rows*6 */
  /* z-slab threading (the MPI build's z partition, general_mesh_subroutines.F90:868-972,
   * as threads): the scan-order counters of scatter and connect at the start
   * of each z-plane, so that planes can run in any order */
  /* host-path faces (thin layer, diode, spice, switch surface materials):
   * per special face code the host face number (1-based, 0 = device face),
   * their count, and the four face voltages per host face that the host
   * wrote for the coming connect */
  int32_t* hf_of_code;
  int n_host_faces;
  double* hf_tot;
  int hf_ready;
  int threads;
  int32_t *pl_out, *pl_exc;   /* output / excitation cell numbers before plane cz (nz+1) */
  int32_t *pl_fexc, *pl_fout; /* excitation / output face numbers before plane cz (nz+1) */
};

/* PML_material_data_type words (PML_module.F90:131-156), without the
 * never-read Ix_m1..Iz_m1. */
enum {
  PM_Ixt_m1, PM_Iyt_m1, PM_Izt_m1, PM_Vcx_m1, PM_Vcy_m1, PM_Vcz_m1, PM_Vcx, PM_Vcy, PM_Vcz,
  PM_Vi_m1, /* 12 words */
  PM_Vxt_m1 = PM_Vi_m1 + 12, PM_Vyt_m1, PM_Vzt_m1,
  PM_Vyxt_m1, PM_Vzxt_m1, PM_Vxyt_m1, PM_Vzyt_m1, PM_Vyzt_m1, PM_Vxzt_m1,
  PM_Vosxr, PM_Vosyr, PM_Voszr, PM_Vosx, PM_Vosy, PM_Vosz, PM_Vosx_m1, PM_Vosy_m1, PM_Vosz_m1,
  PM_Vssxr, PM_Vssyr, PM_Vsszr, PM_Vsx, PM_Vsy, PM_Vsz, PM_Vsx_m1, PM_Vsy_m1, PM_Vsz_m1,
  PM_N
};

/* PML_data_type words (PML_module.F90:117-129). */
enum {
  PS_Vxt, PS_Vyt, PS_Vzt, PS_Ix, PS_Iy, PS_Iz, PS_Ixt, PS_Iyt, PS_Izt,
  PS_Vi, /* 12 words */
  PS_Vyxt = PS_Vi + 12, PS_Vzxt, PS_Vxyt, PS_Vzyt, PS_Vyzt, PS_Vxzt, PS_N
};

static int fail(oracle_handle* h, int code, const char* msg) {
#pragma omp critical(oracle_error)
  snprintf(h->err, sizeof h->err, "%s", msg);
  return code;
}

static inline double* cellV(oracle_handle* h, int cx, int cy, int cz) {
  /* V(1,cx,cy,cz) with cx in 0..nx+1, cy in 0..ny+1, cz in 1..nz */
  size_t i = (size_t)cx + (size_t)(h->nx + 2) * ((size_t)cy + (size_t)(h->ny + 2) * (size_t)(cz - 1));
  return h->V + 12 * i - 1; /* so that p[port] with port 1..12 works */
}

int oracle_create(const tlm_grid* g, int32_t form, oracle_handle** out) {
  if (!g || !out || g->nx < 1 || g->ny < 1 || g->nz < 1) return TLM_ERR_ARG;
  oracle_handle* h = (oracle_handle*)calloc(1, sizeof *h);
  if (!h) return TLM_ERR_NOMEM;
  (void)form; /* the oracle is FP64 throughout */
  h->g = *g;
  h->nx = g->nx; h->ny = g->ny; h->nz = g->nz;
  size_t n = (size_t)12 * (g->nx + 2) * (g->ny + 2) * g->nz;
  h->V = (double*)calloc(n, sizeof(double));
  if (!h->V) { free(h); return TLM_ERR_NOMEM; }
  h->cap = 4096;
  h->threads = 1;
  *out = h;
  return TLM_OK;
}

void oracle_destroy(oracle_handle* h) {
  if (!h) return;
  free(h->V); free(h->cell_code); free(h->c_mat); free(h->c_filt); free(h->c_pml);
  free(h->c_exc); free(h->c_out); free(h->c_cable); free(h->face_code); free(h->f_mat); free(h->f_exc);
  free(h->f_out); free(h->vmat); free(h->smat); free(h->eps_w); free(h->mu_w);
  free(h->eps_ord); free(h->mu_ord); free(h->pml_par); free(h->pml_cell_param);
  free(h->pml_state); free(h->value); free(h->value_face); free(h->ps);
  free(h->pml_mat_pos); free(h->pml_mat_state);
  free(h->cell_exc_field); free(h->cell_exc_type); free(h->face_exc_field);
  free(h->face_exc_type); free(h->huy); free(h->cell_out); free(h->face_out); free(h->fring);
  free(h->pt_row); free(h->pt_comp); free(h->freq); free(h->dft); free(h->ring);
  free(h->fo); free(h->fo_out); free(h->fo_row); free(h->fo_side); free(h->fo_norm); free(h->fo_acc);
  free(h->ff); free(h->ff_row); free(h->ff_side); free(h->ff_norm); free(h->ff_centre); free(h->ff_J); free(h->ff_M);
  for (int i = 0; i < h->n_fd; i++) free(h->fd_frames[i]);
  free(h->fd); free(h->fd_row); free(h->fd_side); free(h->fd_norm); free(h->fd_frames); free(h->fd_nframes);
  free(h->fd_capframes); free(h->fd_first);
  free(h->rcs_row); free(h->rcs_side); free(h->rcs_norm); free(h->rcs_centre); free(h->rcs_E);
  free(h->pl_out); free(h->pl_exc); free(h->pl_fexc); free(h->pl_fout);
  free(h->hf_of_code); free(h->hf_tot);
  free(h);
}

const char* oracle_last_error(const oracle_handle* h) { return h ? h->err : "null handle"; }
int64_t oracle_device_bytes(const oracle_handle* h) { (void)h; return 0; }

static int alloc_filters(oracle_handle* h);

static int32_t* dup_i32(const int32_t* p, size_t n) {
  int32_t* q = (int32_t*)malloc((n ? n : 1) * sizeof(int32_t));
  if (q && p && n) memcpy(q, p, n * sizeof(int32_t));
  else if (q) memset(q, 0, (n ? n : 1) * sizeof(int32_t));
  return q;
}

int oracle_set_cell_codes(oracle_handle* h, const int32_t* code, int32_t n_special,
                          const int32_t* to_material, const int32_t* to_pml,
                          const int32_t* to_excitation, const int32_t* to_output,
                          const int32_t* to_cable) {
  size_t N = (size_t)h->nx * h->ny * h->nz;
  free(h->pml_mat_pos); free(h->pml_mat_state);
  h->pml_mat_pos = NULL; h->pml_mat_state = NULL; h->n_pml_mat = 0;
  for (int32_t i = 0; i < n_special; i++)
    if (to_cable && to_cable[i] &&
        ((to_pml && to_pml[i]) || (to_excitation && to_excitation[i]) || (to_output && to_output[i])))
      return fail(h, TLM_ERR_UNSUPPORTED, "host-path cells carry no PML, excitation or output");
  free(h->pl_out); free(h->pl_exc);
  h->pl_out = h->pl_exc = NULL;
  h->cell_code = dup_i32(code, N);
  h->n_special = n_special;
  h->c_mat = dup_i32(to_material, n_special);
  h->c_filt = dup_i32(to_material ? to_material + n_special : NULL, n_special);
  h->c_pml = dup_i32(to_pml, n_special);
  h->c_exc = dup_i32(to_excitation, n_special);
  h->c_out = dup_i32(to_output, n_special);
  free(h->c_cable);
  h->c_cable = dup_i32(to_cable, n_special);
  int no = 0, ne = 0;
  for (int32_t i = 0; i < n_special; i++) { no += h->c_out[i] != 0; ne += h->c_exc[i] != 0; }
  h->n_out_rows = no;
  h->n_exc_rows = ne;
  free(h->cell_out);
  h->cell_out = (double*)calloc((size_t)(no ? no : 1) * 6, sizeof(double));
  return alloc_filters(h);
}

int oracle_set_face_codes(oracle_handle* h, const int32_t* code, int32_t n_special,
                          const int32_t* to_material, const int32_t* to_excitation,
                          const int32_t* to_output, const int32_t* to_cable) {
  int nx = h->nx, ny = h->ny, nz = h->nz;
  size_t nf = (size_t)(nx - 1) * ny * nz + (size_t)nx * (ny - 1) * nz + (size_t)nx * ny * (nz - 1);
  /* cable faces (to_cable != 0): connected as any special face; the cable
     junction itself (connect_normal_to_x.F90:317-319, face_cable_junction)
     touches only the cable model and runs on the host after the connect */
  (void)to_cable;
  free(h->pl_fexc); free(h->pl_fout);
  h->pl_fexc = h->pl_fout = NULL;
  free(h->hf_of_code); h->hf_of_code = NULL;
  h->n_faces = (int)nf;
  h->face_code = dup_i32(code, nf);
  h->n_special_faces = n_special;
  h->f_mat = dup_i32(to_material, n_special);
  h->f_exc = dup_i32(to_excitation, n_special);
  h->f_out = dup_i32(to_output, n_special);
  int no = 0;
  for (int32_t i = 0; i < n_special; i++) no += h->f_out[i] != 0;
  h->n_face_out_rows = no;
  free(h->face_out);
  free(h->fring); h->fring = NULL; h->fring_count = 0;
  h->face_out = (double*)calloc((size_t)(no ? no : 1) * 12, sizeof(double));
  return TLM_OK;
}

/* Filter storage follows allocate_volume_material_filter_data
 * (TLM_volume_materials.F90:401-518): three filters per dispersive cell,
 * each with w(0:order) zeroed (allocate_Zfilter_response,
 * FILTER/filter_functions.F90:261-286). */
static int alloc_filters(oracle_handle* h) {
  if (!h->cell_code || !h->vmat) return TLM_OK;
  int maxf = 0;
  for (int i = 0; i < h->n_special; i++)
    if (h->c_mat[i] && h->c_filt[i] + 2 > maxf) maxf = h->c_filt[i] + 2;
  h->n_filters = maxf;
  free(h->eps_w); free(h->mu_w); free(h->eps_ord); free(h->mu_ord);
  h->eps_w = (double*)calloc((size_t)(maxf + 1) * NW, sizeof(double));
  h->mu_w = (double*)calloc((size_t)(maxf + 1) * NW, sizeof(double));
  h->eps_ord = (int32_t*)calloc((size_t)(maxf + 1), sizeof(int32_t));
  h->mu_ord = (int32_t*)calloc((size_t)(maxf + 1), sizeof(int32_t));
  for (int i = 0; i < h->n_special; i++) {
    int m = h->c_mat[i];
    if (!m) continue;
    if (m < 1 || m > h->n_vmat) return fail(h, TLM_ERR_ARG, "material number out of range");
    const tlm_volume_material* vm = &h->vmat[m - 1];
    if (vm->type != TLM_VMAT_DISPERSIVE) continue;
    int f = h->c_filt[i];
    if (f < 1) return fail(h, TLM_ERR_ARG, "dispersive cell without filter number");
    for (int k = 0; k < 3; k++) {
      h->eps_ord[f + k] = vm->eps_filter_exists ? vm->eps_order : 0;
      h->mu_ord[f + k] = vm->mu_filter_exists ? vm->mu_order : 0;
    }
  }
  return TLM_OK;
}

int oracle_set_volume_materials(oracle_handle* h, int32_t n, const tlm_volume_material* m) {
  free(h->vmat);
  free(h->pml_mat_pos); free(h->pml_mat_state);
  h->pml_mat_pos = NULL; h->pml_mat_state = NULL; h->n_pml_mat = 0;
  h->n_vmat = n;
  h->vmat = (tlm_volume_material*)malloc((size_t)(n ? n : 1) * sizeof *m);
  if (n) memcpy(h->vmat, m, (size_t)n * sizeof *m);
  for (int i = 0; i < n; i++) {
    if (m[i].eps_order < 0 || m[i].eps_order > TLM_MAX_FILTER_ORDER ||
        m[i].mu_order < 0 || m[i].mu_order > TLM_MAX_FILTER_ORDER)
      return fail(h, TLM_ERR_ARG, "filter order out of range");
  }
  return alloc_filters(h);
}

/* Surface material types whose connect stays on the host
 * (connect_normal_to_x.F90:75-283). */
static int host_smat(int type) {
  return type == TLM_SMAT_DISPERSIVE || type == TLM_SMAT_ANISOTROPIC_DISPERSIVE || type == TLM_SMAT_DIODE ||
         type == TLM_SMAT_SPICE || type == TLM_SMAT_SWITCH;
}

int oracle_set_surface_materials(oracle_handle* h, int32_t n, const int32_t* type) {
  for (int i = 0; i < n; i++)
    if (type[i] != TLM_SMAT_PEC && type[i] != TLM_SMAT_PMC && type[i] != TLM_SMAT_FREE_SPACE && !host_smat(type[i]))
      return fail(h, TLM_ERR_UNSUPPORTED, "unknown surface material type");
  free(h->hf_of_code); h->hf_of_code = NULL;
  free(h->smat);
  h->n_smat = n;
  h->smat = dup_i32(type, n);
  return TLM_OK;
}

int oracle_set_pml(oracle_handle* h, int32_t n_params, const tlm_pml_parameter* p,
                   int32_t n_cells, const int32_t* cell_param) {
  free(h->pml_par); free(h->pml_cell_param); free(h->pml_state);
  free(h->pml_mat_pos); free(h->pml_mat_state);
  h->pml_mat_pos = NULL; h->pml_mat_state = NULL; h->n_pml_mat = 0;
  h->n_pml_params = n_params;
  h->pml_par = (tlm_pml_parameter*)malloc((size_t)(n_params ? n_params : 1) * sizeof *p);
  if (n_params) memcpy(h->pml_par, p, (size_t)n_params * sizeof *p);
  h->n_pml_cells = n_cells;
  h->pml_cell_param = dup_i32(cell_param, n_cells);
  h->pml_state = (double*)calloc((size_t)(n_cells ? n_cells : 1) * PS_N, sizeof(double));
  return TLM_OK;
}

int oracle_set_excitation_functions(oracle_handle* h, int32_t n_functions,
                                    const double* value, const double* value_face) {
  size_t n = (size_t)n_functions * h->g.n_timesteps;
  free(h->value); free(h->value_face);
  h->n_functions = n_functions;
  h->value = (double*)malloc((n ? n : 1) * sizeof(double));
  h->value_face = (double*)malloc((n ? n : 1) * sizeof(double));
  if (n) { memcpy(h->value, value, n * sizeof(double)); memcpy(h->value_face, value_face, n * sizeof(double)); }
  return TLM_OK;
}

int oracle_set_point_sources(oracle_handle* h, int32_t n, const tlm_point_source* s,
                             int32_t n_cell_rows, int32_t n_face_rows) {
  free(h->ps); free(h->cell_exc_field); free(h->cell_exc_type);
  free(h->face_exc_field); free(h->face_exc_type);
  h->n_ps = n;
  h->ps = (tlm_point_source*)malloc((size_t)(n ? n : 1) * sizeof *s);
  if (n) memcpy(h->ps, s, (size_t)n * sizeof *s);
  h->n_cell_rows = n_cell_rows;
  h->n_face_rows = n_face_rows;
  h->cell_exc_field = (double*)calloc((size_t)(n_cell_rows ? n_cell_rows : 1) * 9, sizeof(double));
  h->cell_exc_type = (int32_t*)malloc((size_t)(n_cell_rows ? n_cell_rows : 1) * 6 * sizeof(int32_t));
  h->face_exc_field = (double*)calloc((size_t)(n_face_rows ? n_face_rows : 1) * 24, sizeof(double));
  h->face_exc_type = (int32_t*)malloc((size_t)(n_face_rows ? n_face_rows : 1) * 12 * sizeof(int32_t));
  /* set_cell_update_codes / set_face_update_codes initialise the types soft */
  for (int i = 0; i < (n_cell_rows ? n_cell_rows : 1) * 6; i++) h->cell_exc_type[i] = 1;
  for (int i = 0; i < (n_face_rows ? n_face_rows : 1) * 12; i++) h->face_exc_type[i] = 1;
  for (int i = 0; i < n; i++) {
    if (s[i].function < 1 || s[i].function > h->n_functions) return fail(h, TLM_ERR_ARG, "point source function out of range");
    if (s[i].component < 1 || s[i].component > (s[i].at_face ? 12 : 9)) return fail(h, TLM_ERR_ARG, "point source component out of range");
    int rows = s[i].at_face ? n_face_rows : n_cell_rows;
    if (s[i].row < 1 || s[i].row > rows) return fail(h, TLM_ERR_ARG, "point source row out of range");
  }
  return TLM_OK;
}

int oracle_set_huygens_surface(oracle_handle* h, int32_t function, double offset_min,
                               const double* Ei, const double* Hi, int32_t n_faces,
                               const tlm_huygens_face* faces) {
  free(h->huy);
  h->huy_function = function;
  h->huy_offset_min = offset_min;
  for (int i = 0; i < 3; i++) { h->Ei[i] = Ei[i]; h->Hi[i] = Hi[i]; }
  h->n_huy = n_faces;
  h->huy = (tlm_huygens_face*)malloc((size_t)(n_faces ? n_faces : 1) * sizeof *faces);
  if (n_faces) memcpy(h->huy, faces, (size_t)n_faces * sizeof *faces);
  if (n_faces && (function < 1 || function > h->n_functions)) return fail(h, TLM_ERR_ARG, "Huygens function out of range");
  for (int i = 0; i < n_faces; i++)
    if (faces[i].row < 0 || faces[i].row > h->n_face_rows) return fail(h, TLM_ERR_ARG, "Huygens face row out of range");
  return TLM_OK;
}

int oracle_set_output_points(oracle_handle* h, int32_t n, const int32_t* row, const int32_t* comp) {
  free(h->pt_row); free(h->pt_comp); free(h->dft);
  h->n_points = n;
  h->pt_row = dup_i32(row, n);
  h->pt_comp = dup_i32(comp, n);
  for (int i = 0; i < n; i++)
    if (row[i] < 1 || row[i] > h->n_out_rows || comp[i] < 1 || comp[i] > 6)
      return fail(h, TLM_ERR_ARG, "output point row/component out of range");
  h->dft = (double*)calloc((size_t)(n ? n : 1) * (h->n_freq ? h->n_freq : 1) * 2, sizeof(double));
  return TLM_OK;
}

int oracle_set_dft_frequencies(oracle_handle* h, int32_t n, const double* f) {
  free(h->freq); free(h->dft);
  h->n_freq = n;
  h->freq = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
  if (n) memcpy(h->freq, f, (size_t)n * sizeof(double));
  h->dft = (double*)calloc((size_t)(h->n_points ? h->n_points : 1) * (n ? n : 1) * 2, sizeof(double));
  return TLM_OK;
}

int oracle_upload_links(oracle_handle* h, const double* V) {
  size_t n = (size_t)12 * (h->nx + 2) * (h->ny + 2) * h->nz;
  memcpy(h->V, V, n * sizeof(double));
  return TLM_OK;
}

int oracle_download_links(oracle_handle* h, double* V) {
  for (int cz = 1; cz <= h->nz; cz++)
    for (int cy = 1; cy <= h->ny; cy++)
      for (int cx = 1; cx <= h->nx; cx++) {
        size_t i = (size_t)cx + (size_t)(h->nx + 2) * ((size_t)cy + (size_t)(h->ny + 2) * (size_t)(cz - 1));
        memcpy(V + 12 * i, h->V + 12 * i, 12 * sizeof(double));
      }
  return TLM_OK;
}

/* The oracle holds no device memory: the allocation-free download is the
 * download. */
int oracle_download_links_noalloc(oracle_handle* h, double* V) { return oracle_download_links(h, V); }

/* ---- excitation ---------------------------------------------------------- */

/* get_interpolated_excitation_value, EXCITATION/excitation.F90:645-690.
 * NINT rounds half away from zero, as C round() does. */
static double interpolated_value(oracle_handle* h, int timestep, double offset, int function) {
  double dt = h->g.dt;
  int nts = h->g.n_timesteps;
  double t_offset = (offset - h->huy_offset_min) / (C0 / h->g.reduced_c_factor);
  double local_time = ((double)(timestep - 1)) * dt - t_offset;
  int timestep1 = (int)round(local_time / dt);
  int timestep2 = timestep1 + 1;
  double delta_t = local_time - (double)timestep1 * dt;
  const double* vf = h->value_face + (size_t)(function - 1) * nts;
  double value1 = (timestep1 >= 1 && timestep1 <= nts) ? vf[timestep1 - 1] : 0.0;
  double value2 = (timestep2 >= 1 && timestep2 <= nts) ? vf[timestep2 - 1] : 0.0;
  return value1 + delta_t * (value2 - value1) / dt;
}

int oracle_excitation(oracle_handle* h, int32_t timestep) {
  if (timestep < 1 || timestep > h->g.n_timesteps) return fail(h, TLM_ERR_ARG, "timestep out of range");
  int nts = h->g.n_timesteps;
  /* excitation.F90:611-612: reset all excitation fields */
  if (h->cell_exc_field) memset(h->cell_exc_field, 0, (size_t)(h->n_cell_rows ? h->n_cell_rows : 1) * 9 * sizeof(double));
  if (h->face_exc_field) memset(h->face_exc_field, 0, (size_t)(h->n_face_rows ? h->n_face_rows : 1) * 24 * sizeof(double));
  /* Point_excitation */
  for (int i = 0; i < h->n_ps; i++) {
    const tlm_point_source* s = &h->ps[i];
    int f = s->function, c = s->component, r = s->row;
    if (!s->at_face) {
      double value = h->value[(size_t)(f - 1) * nts + (timestep - 1)];
      h->cell_exc_field[(size_t)(r - 1) * 9 + (c - 1)] = h->cell_exc_field[(size_t)(r - 1) * 9 + (c - 1)] + value;
      if (c <= 6) h->cell_exc_type[(size_t)(r - 1) * 6 + (c - 1)] = s->source_type;
    } else {
      double value = h->value_face[(size_t)(f - 1) * nts + (timestep - 1)];
      for (int side = 0; side < 2; side++) {
        double* fe = &h->face_exc_field[((size_t)(r - 1) * 2 + side) * 12];
        fe[c - 1] = fe[c - 1] + value;
        if (c <= 6) h->face_exc_type[((size_t)(r - 1) * 2 + side) * 6 + (c - 1)] = s->source_type;
      }
    }
  }
  /* Huygens_surface_excitation, huygens_surface.F90:760-842 */
  for (int i = 0; i < h->n_huy; i++) {
    const tlm_huygens_face* hf = &h->huy[i];
    double value = interpolated_value(h, timestep, hf->offset, h->huy_function);
    double normx = (double)hf->nx, normy = (double)hf->ny, normz = (double)hf->nz;
    double Js[3], Ms[3];
    Js[0] = (normy * value * (h->Hi[2]) - normz * value * (h->Hi[1]));
    Js[1] = (normz * value * (h->Hi[0]) - normx * value * (h->Hi[2]));
    Js[2] = (normx * value * (h->Hi[1]) - normy * value * (h->Hi[0]));
    Ms[0] = -(normy * value * (h->Ei[2]) - normz * value * (h->Ei[1]));
    Ms[1] = -(normz * value * (h->Ei[0]) - normx * value * (h->Ei[2]));
    Ms[2] = -(normx * value * (h->Ei[1]) - normy * value * (h->Ei[0]));
    if (hf->row == 0) continue; /* face not on an interior connect face */
    for (int side = 0; side < 2; side++) {
      double* fe = &h->face_exc_field[((size_t)(hf->row - 1) * 2 + side) * 12];
      for (int k = 0; k < 3; k++) fe[Jx - 1 + k] = fe[Jx - 1 + k] + Js[k];
      for (int k = 0; k < 3; k++) fe[Mx - 1 + k] = fe[Mx - 1 + k] + Ms[k];
    }
  }
  return TLM_OK;
}

/* ---- scatter ------------------------------------------------------------- */

/* evaluate_Zfilter, FILTER/filter_subroutines.F90:510-540 */
static double evaluate_Zfilter(const double* a, const double* b, double* w, int order, double G) {
  double W0 = G;
  for (int i = 1; i <= order; i++) W0 = W0 - b[i] * w[i];
  W0 = W0 / b[0];
  w[0] = W0;
  double F0 = 0.0;
  for (int i = 0; i <= order; i++) F0 = F0 + a[i] * w[i];
  return F0;
}

/* timeshift_Zfilter, filter_subroutines.F90:557-575 */
static void timeshift_Zfilter(double* w, int order) {
  for (int i = order; i >= 1; i--) w[i] = w[i - 1];
}

/* add_shunt_admittance / add_series_impedance, scatter.F90:606-661 */
static void add_shunt(double* V1, double* Y1, double V2, double Y2) {
  double V = (*V1 * *Y1 + V2 * Y2) / (*Y1 + Y2);
  double Y = *Y1 + Y2;
  *V1 = V;
  *Y1 = Y;
}
static void add_series(double* V1, double* Z1, double V2, double Z2) {
  double V = *V1 + V2;
  double Z = *Z1 + Z2;
  *V1 = V;
  *Z1 = Z;
}

/* The 12-port scatter, scatter.F90:240-272 (and :517-547). */
static inline void scatter_ports(double* v, double Vx, double Vy, double Vz,
                                 double Ix, double Iy, double Iz) {
  double V_min, V_max;
  V_min = v[Vx_ymin]; V_max = v[Vx_ymax];
  v[Vx_ymin] = Vx - Z0 * Iz - V_max;
  v[Vx_ymax] = Vx + Z0 * Iz - V_min;
  V_min = v[Vx_zmin]; V_max = v[Vx_zmax];
  v[Vx_zmin] = Vx + Z0 * Iy - V_max;
  v[Vx_zmax] = Vx - Z0 * Iy - V_min;
  V_min = v[Vy_zmin]; V_max = v[Vy_zmax];
  v[Vy_zmin] = Vy - Z0 * Ix - V_max;
  v[Vy_zmax] = Vy + Z0 * Ix - V_min;
  V_min = v[Vy_xmin]; V_max = v[Vy_xmax];
  v[Vy_xmin] = Vy + Z0 * Iz - V_max;
  v[Vy_xmax] = Vy - Z0 * Iz - V_min;
  V_min = v[Vz_xmin]; V_max = v[Vz_xmax];
  v[Vz_xmin] = Vz - Z0 * Iy - V_max;
  v[Vz_xmax] = Vz + Z0 * Iy - V_min;
  V_min = v[Vz_ymin]; V_max = v[Vz_ymax];
  v[Vz_ymin] = Vz + Z0 * Ix - V_max;
  v[Vz_ymax] = Vz - Z0 * Ix - V_min;
}

/* PML port aliases, PML_module.F90:60-71 */
#define Vxny 1
#define Vxnz 2
#define Vxpy 3
#define Vxpz 4
#define Vynx 5
#define Vynz 6
#define Vypx 7
#define Vypz 8
#define Vznx 9
#define Vzny 10
#define Vzpx 11
#define Vzpy 12

/* scatter_PML.F90:22-313, the free-space PML cell update. */
static void scatter_pml_cell(oracle_handle* h, double* v, int PML_cell, int output_number,
                             int* output_cell_number) {
  const double dl = h->g.dl;
  /* STAGE 1 */
  const tlm_pml_parameter* P = &h->pml_par[h->pml_cell_param[PML_cell - 1] - 1];
  double ax = P->ax, ay = P->ay, az = P->az;
  double exp_xi = P->exp_x, exp_yi = P->exp_y, exp_zi = P->exp_z;
  double exp_xr = P->exp_x, exp_yr = P->exp_y, exp_zr = P->exp_z;
  double Csx = P->Csx, Csy = P->Csy, Csz = P->Csz;
  double Csx2 = P->Csx2, Csy2 = P->Csy2, Csz2 = P->Csz2;
  double Csy_sz = P->Csy_sz, Csz_sx = P->Csz_sx, Csx_sy = P->Csx_sy;
  double* S = &h->pml_state[(size_t)(PML_cell - 1) * PS_N];
  /* STAGE 2 */
  double last_Ix = S[PS_Ix], last_Iy = S[PS_Iy], last_Iz = S[PS_Iz];
  double last_Vi[13];
  for (int i = 1; i <= 12; i++) last_Vi[i] = S[PS_Vi + i - 1];
  double last_Vxt = S[PS_Vxt], last_Vyt = S[PS_Vyt], last_Vzt = S[PS_Vzt];
  double last_Vyxt = S[PS_Vyxt], last_Vzxt = S[PS_Vzxt], last_Vxyt = S[PS_Vxyt];
  double last_Vzyt = S[PS_Vzyt], last_Vyzt = S[PS_Vyzt], last_Vxzt = S[PS_Vxzt];
  /* STAGE 3 */
  v[Vynx] = v[Vynx] * exp_yi; v[Vypx] = v[Vypx] * exp_yi;
  v[Vznx] = v[Vznx] * exp_zi; v[Vzpx] = v[Vzpx] * exp_zi;
  v[Vxny] = v[Vxny] * exp_xi; v[Vxpy] = v[Vxpy] * exp_xi;
  v[Vzny] = v[Vzny] * exp_zi; v[Vzpy] = v[Vzpy] * exp_zi;
  v[Vxnz] = v[Vxnz] * exp_xi; v[Vxpz] = v[Vxpz] * exp_xi;
  v[Vynz] = v[Vynz] * exp_yi; v[Vypz] = v[Vypz] * exp_yi;
  /* STAGE 4 */
  double Vx = (v[Vynx] + v[Vypx] + v[Vznx] + v[Vzpx]) / 2.0;
  double Vy = (v[Vxny] + v[Vxpy] + v[Vzny] + v[Vzpy]) / 2.0;
  double Vz = (v[Vxnz] + v[Vxpz] + v[Vynz] + v[Vypz]) / 2.0;
  double Ix = (v[Vypz] - v[Vzpy] - v[Vynz] + v[Vzny]) / (2.0 * Z0);
  double Iy = (v[Vzpx] - v[Vxpz] - v[Vznx] + v[Vxnz]) / (2.0 * Z0);
  double Iz = (v[Vxpy] - v[Vypx] - v[Vxny] + v[Vynx]) / (2.0 * Z0);
  /* STAGE 5 */
  double Vx_PML_shunt = ax * (-(last_Vi[Vznx] + last_Vi[Vzpx] + last_Vi[Vynx] + last_Vi[Vypx]) / 2.0
                              + Csy * ((v[Vznx] + v[Vzpx]) + (last_Vi[Vznx] + last_Vi[Vzpx]))
                              + Csz * ((v[Vynx] + v[Vypx]) + (last_Vi[Vynx] + last_Vi[Vypx]))
                              + Csy_sz * last_Vxt);
  double Vy_PML_shunt = ay * (-(last_Vi[Vxny] + last_Vi[Vxpy] + last_Vi[Vzny] + last_Vi[Vzpy]) / 2.0
                              + Csz * ((v[Vxny] + v[Vxpy]) + (last_Vi[Vxny] + last_Vi[Vxpy]))
                              + Csx * ((v[Vzny] + v[Vzpy]) + (last_Vi[Vzny] + last_Vi[Vzpy]))
                              + Csz_sx * last_Vyt);
  double Vz_PML_shunt = az * (-(last_Vi[Vynz] + last_Vi[Vypz] + last_Vi[Vxnz] + last_Vi[Vxpz]) / 2.0
                              + Csx * ((v[Vynz] + v[Vypz]) + (last_Vi[Vynz] + last_Vi[Vypz]))
                              + Csy * ((v[Vxnz] + v[Vxpz]) + (last_Vi[Vxnz] + last_Vi[Vxpz]))
                              + Csx_sy * last_Vzt);
  /* STAGE 6 */
  double Vxt = ax * Vx + Vx_PML_shunt;
  double Vyt = ay * Vy + Vy_PML_shunt;
  double Vzt = az * Vz + Vz_PML_shunt;
  /* STAGE 7a */
  double Ixt = 0.0, Iyt = 0.0, Izt = 0.0;
  if (output_number != 0) {
    double last_Ixt = S[PS_Ixt], last_Iyt = S[PS_Iyt], last_Izt = S[PS_Izt];
    Ixt = ax * (Ix - last_Ix + Csx2 * (Ix + last_Ix) + Csy_sz * last_Ixt);
    Iyt = ay * (Iy - last_Iy + Csy2 * (Iy + last_Iy) + Csz_sx * last_Iyt);
    Izt = az * (Iz - last_Iz + Csz2 * (Iz + last_Iz) + Csx_sy * last_Izt);
  }
  /* STAGE 7 */
  double Vxy_PML_series = az * (Z0 * (-last_Iz + Csy2 * (Iz + last_Iz)) + Csx_sy * last_Vxyt);
  double Vyx_PML_series = az * (Z0 * (-last_Iz + Csx2 * (Iz + last_Iz)) + Csx_sy * last_Vyxt);
  double Vzx_PML_series = ay * (Z0 * (-last_Iy + Csx2 * (Iy + last_Iy)) + Csz_sx * last_Vzxt);
  double Vxz_PML_series = ay * (Z0 * (-last_Iy + Csz2 * (Iy + last_Iy)) + Csz_sx * last_Vxzt);
  double Vyz_PML_series = ax * (Z0 * (-last_Ix + Csz2 * (Ix + last_Ix)) + Csy_sz * last_Vyzt);
  double Vzy_PML_series = ax * (Z0 * (-last_Ix + Csy2 * (Ix + last_Ix)) + Csy_sz * last_Vzyt);
  /* STAGE 8 */
  double Vxyt = az * Iz * Z0 + Vxy_PML_series;
  double Vyxt = az * Iz * Z0 + Vyx_PML_series;
  double Vzxt = ay * Iy * Z0 + Vzx_PML_series;
  double Vxzt = ay * Iy * Z0 + Vxz_PML_series;
  double Vyzt = ax * Ix * Z0 + Vyz_PML_series;
  double Vzyt = ax * Ix * Z0 + Vzy_PML_series;
  /* STAGE 9 */
  S[PS_Ix] = Ix; S[PS_Iy] = Iy; S[PS_Iz] = Iz;
  for (int i = 1; i <= 12; i++) S[PS_Vi + i - 1] = v[i];
  S[PS_Vxt] = Vxt; S[PS_Vyt] = Vyt; S[PS_Vzt] = Vzt;
  if (output_number != 0) { S[PS_Ixt] = Ixt; S[PS_Iyt] = Iyt; S[PS_Izt] = Izt; }
  S[PS_Vxyt] = Vxyt; S[PS_Vyxt] = Vyxt;
  S[PS_Vzxt] = Vzxt; S[PS_Vxzt] = Vxzt;
  S[PS_Vzyt] = Vzyt; S[PS_Vyzt] = Vyzt;
  /* STAGE 10 */
  double Vynx_PML = Vx_PML_shunt - Vyx_PML_series;
  double Vypx_PML = Vx_PML_shunt + Vyx_PML_series;
  double Vznx_PML = Vx_PML_shunt + Vzx_PML_series;
  double Vzpx_PML = Vx_PML_shunt - Vzx_PML_series;
  double Vxny_PML = Vy_PML_shunt + Vxy_PML_series;
  double Vxpy_PML = Vy_PML_shunt - Vxy_PML_series;
  double Vzny_PML = Vy_PML_shunt - Vzy_PML_series;
  double Vzpy_PML = Vy_PML_shunt + Vzy_PML_series;
  double Vynz_PML = Vz_PML_shunt + Vyz_PML_series;
  double Vypz_PML = Vz_PML_shunt - Vyz_PML_series;
  double Vxnz_PML = Vz_PML_shunt - Vxz_PML_series;
  double Vxpz_PML = Vz_PML_shunt + Vxz_PML_series;
  /* STAGE 11 */
  double V_max, V_min;
  V_max = v[Vypx]; V_min = v[Vynx];
  v[Vynx] = ax * Vx - az * Z0 * Iz - V_max + Vynx_PML;
  v[Vypx] = ax * Vx + az * Z0 * Iz - V_min + Vypx_PML;
  V_max = v[Vzpx]; V_min = v[Vznx];
  v[Vznx] = ax * Vx + ay * Z0 * Iy - V_max + Vznx_PML;
  v[Vzpx] = ax * Vx - ay * Z0 * Iy - V_min + Vzpx_PML;
  V_max = v[Vzpy]; V_min = v[Vzny];
  v[Vzny] = ay * Vy - ax * Z0 * Ix - V_max + Vzny_PML;
  v[Vzpy] = ay * Vy + ax * Z0 * Ix - V_min + Vzpy_PML;
  V_max = v[Vxpy]; V_min = v[Vxny];
  v[Vxny] = ay * Vy + az * Z0 * Iz - V_max + Vxny_PML;
  v[Vxpy] = ay * Vy - az * Z0 * Iz - V_min + Vxpy_PML;
  V_max = v[Vxpz]; V_min = v[Vxnz];
  v[Vxnz] = az * Vz - ay * Z0 * Iy - V_max + Vxnz_PML;
  v[Vxpz] = az * Vz + ay * Z0 * Iy - V_min + Vxpz_PML;
  V_max = v[Vypz]; V_min = v[Vynz];
  v[Vynz] = az * Vz + ax * Z0 * Ix - V_max + Vynz_PML;
  v[Vypz] = az * Vz - ax * Z0 * Ix - V_min + Vypz_PML;
  /* STAGE 12 */
  if (output_number != 0) {
    *output_cell_number += 1;
    double* fo = &h->cell_out[(size_t)(*output_cell_number - 1) * 6];
    fo[0] = -Vxt / dl; fo[1] = -Vyt / dl; fo[2] = -Vzt / dl;
    fo[3] = Ixt / dl; fo[4] = Iyt / dl; fo[5] = Izt / dl;
  }
  /* STAGE 13 */
  v[Vynx] = v[Vynx] * exp_yr; v[Vypx] = v[Vypx] * exp_yr;
  v[Vznx] = v[Vznx] * exp_zr; v[Vzpx] = v[Vzpx] * exp_zr;
  v[Vxny] = v[Vxny] * exp_xr; v[Vxpy] = v[Vxpy] * exp_xr;
  v[Vzny] = v[Vzny] * exp_zr; v[Vzpy] = v[Vzpy] * exp_zr;
  v[Vxnz] = v[Vxnz] * exp_xr; v[Vxpz] = v[Vxpz] * exp_xr;
  v[Vynz] = v[Vynz] * exp_yr; v[Vypz] = v[Vypz] * exp_yr;
}

/* allocate_PML_material_data (TLM_PML_volume_materials.F90:597-746):
 * number the PML cells holding a DISPERSIVE-type material in scan order,
 * zero their state. */
static int setup_pml_materials(oracle_handle* h) {
  if (h->pml_mat_pos) return TLM_OK;
  h->pml_mat_pos = (int32_t*)calloc((size_t)(h->n_pml_cells ? h->n_pml_cells : 1), sizeof(int32_t));
  long long N = (long long)h->nx * h->ny * h->nz;
  int n = 0;
  for (long long c = 0; c < N; c++) {
    int code = h->cell_code[c];
    if (!code) continue;
    int PML_cell = h->c_pml[code - 1];
    if (!PML_cell) continue;
    int m = h->c_mat[code - 1];
    if (m && h->vmat[m - 1].type == TLM_VMAT_DISPERSIVE) {
      /* set_update_codes.F90:300-333: only order-0 (non-dispersive) S-filters
       * are allowed in the PML; their stub Z-filters are of order <= 1 */
      if (h->vmat[m - 1].eps_order > 1 || h->vmat[m - 1].mu_order > 1)
        return fail(h, TLM_ERR_ARG, "PML fatal error: dispersive material found");
      h->pml_mat_pos[PML_cell - 1] = ++n;
    }
  }
  h->n_pml_mat = n;
  h->pml_mat_state = (double*)calloc((size_t)(n ? n : 1) * PM_N, sizeof(double));
  return TLM_OK;
}

/* scatter_PML_material.F90:22-534, a PML cell holding a non-dispersive
 * material.  The two-step recursion coefficients (STAGE 1) depend only on
 * the PML parameter row and the material; they are recomputed here per cell
 * as the reference does.  PML_cell_data%Ix..Iz, which this update stores
 * from stale scatter locals and nothing reads, are not kept. */
static void scatter_pml_material_cell(oracle_handle* h, double* v, int PML_cell, int material_number,
                                      int output_number, int* output_cell_number) {
  const double dl = h->g.dl, dt = h->g.dt;
  const tlm_volume_material* vm = &h->vmat[material_number - 1];
  /* STAGE 1 */
  double Ysx = vm->Ys_eps_f / Y0, Ysy = vm->Ys_eps_f / Y0, Ysz = vm->Ys_eps_f / Y0;
  double Gex = vm->Ge / Y0, Gey = vm->Ge / Y0, Gez = vm->Ge / Y0;
  double Zsx = vm->Zs_mu_f, Zsy = vm->Zs_mu_f, Zsz = vm->Zs_mu_f;
  double Rmx = vm->Rm, Rmy = vm->Rm, Rmz = vm->Rm;
  const tlm_pml_parameter* P = &h->pml_par[h->pml_cell_param[PML_cell - 1] - 1];
  double sx = P->sx, sy = P->sy, sz = P->sz;
  double exp_xi = P->exp_x, exp_yi = P->exp_y, exp_zi = P->exp_z;
  double exp_xr = P->exp_x, exp_yr = P->exp_y, exp_zr = P->exp_z;
  double GpY2;
  GpY2 = (Gex + Ysx) / 2.0;
  double Ax0 = 8.0 + 4.0 * GpY2 + 2.0 * dt * (sy + sz) * (1.0 + GpY2) + dt * dt * sy * sz * GpY2;
  double Ax1 = -2.0 * (8.0 + 4.0 * GpY2) + 2.0 * (dt * dt * sy * sz * GpY2);
  double Ax2 = 8.0 + 4.0 * GpY2 - 2.0 * dt * (sy + sz) * (1.0 + GpY2) + dt * dt * sy * sz * GpY2;
  GpY2 = (Gey + Ysy) / 2.0;
  double Ay0 = 8.0 + 4.0 * GpY2 + 2.0 * dt * (sz + sx) * (1.0 + GpY2) + dt * dt * sz * sx * GpY2;
  double Ay1 = -2.0 * (8.0 + 4.0 * GpY2) + 2.0 * (dt * dt * sz * sx * GpY2);
  double Ay2 = 8.0 + 4.0 * GpY2 - 2.0 * dt * (sz + sx) * (1.0 + GpY2) + dt * dt * sz * sx * GpY2;
  GpY2 = (Gez + Ysz) / 2.0;
  double Az0 = 8.0 + 4.0 * GpY2 + 2.0 * dt * (sx + sy) * (1.0 + GpY2) + dt * dt * sx * sy * GpY2;
  double Az1 = -2.0 * (8.0 + 4.0 * GpY2) + 2.0 * (dt * dt * sx * sy * GpY2);
  double Az2 = 8.0 + 4.0 * GpY2 - 2.0 * dt * (sx + sy) * (1.0 + GpY2) + dt * dt * sx * sy * GpY2;
  double Bx0 = 4.0 + 2.0 * dt * sy, Bx1 = -8.0, Bx2 = 4.0 - 2.0 * dt * sy;
  double By0 = 4.0 + 2.0 * dt * sz, By1 = -8.0, By2 = 4.0 - 2.0 * dt * sz;
  double Bz0 = 4.0 + 2.0 * dt * sx, Bz1 = -8.0, Bz2 = 4.0 - 2.0 * dt * sx;
  double Cx0 = 4.0 + 2.0 * dt * sz, Cx1 = -8.0, Cx2 = 4.0 - 2.0 * dt * sz;
  double Cy0 = 4.0 + 2.0 * dt * sx, Cy1 = -8.0, Cy2 = 4.0 - 2.0 * dt * sx;
  double Cz0 = 4.0 + 2.0 * dt * sy, Cz1 = -8.0, Cz2 = 4.0 - 2.0 * dt * sy;
  double Dx0 = 4.0 * Ysx + 2.0 * dt * (sy + sz) * Ysx + 1.0 * (dt * dt * sy * sz * Ysx);
  double Dx1 = -8.0 * Ysx + 2.0 * (dt * dt * sy * sz * Ysx);
  double Dx2 = 4.0 * Ysx - 2.0 * dt * (sy + sz) * Ysx + 1.0 * (dt * dt * sy * sz * Ysx);
  double Dy0 = 4.0 * Ysy + 2.0 * dt * (sz + sx) * Ysy + 1.0 * (dt * dt * sz * sx * Ysy);
  double Dy1 = -8.0 * Ysy + 2.0 * (dt * dt * sz * sx * Ysy);
  double Dy2 = 4.0 * Ysy - 2.0 * dt * (sz + sx) * Ysy + 1.0 * (dt * dt * sz * sx * Ysy);
  double Dz0 = 4.0 * Ysz + 2.0 * dt * (sx + sy) * Ysz + 1.0 * (dt * dt * sx * sy * Ysz);
  double Dz1 = -8.0 * Ysz + 2.0 * (dt * dt * sx * sy * Ysz);
  double Dz2 = 4.0 * Ysz - 2.0 * dt * (sx + sy) * Ysz + 1.0 * (dt * dt * sx * sy * Ysz);
  double ZpR;
  ZpR = Zsx + Rmx;
  double Sx0 = 16.0 * Z0 + 4.0 * ZpR + 2.0 * dt * (sy + sz) * (2.0 * Z0 + ZpR) + dt * dt * sy * sz * ZpR;
  double Sx1 = -32.0 * Z0 - 8.0 * ZpR + 2.0 * dt * dt * sy * sz * ZpR;
  double Sx2 = 16.0 * Z0 + 4.0 * ZpR - 2.0 * dt * (sy + sz) * (2.0 * Z0 + ZpR) + dt * dt * sy * sz * ZpR;
  ZpR = Zsy + Rmy;
  double Sy0 = 16.0 * Z0 + 4.0 * ZpR + 2.0 * dt * (sz + sx) * (2.0 * Z0 + ZpR) + dt * dt * sz * sx * ZpR;
  double Sy1 = -32.0 * Z0 - 8.0 * ZpR + 2.0 * dt * dt * sz * sx * ZpR;
  double Sy2 = 16.0 * Z0 + 4.0 * ZpR - 2.0 * dt * (sz + sx) * (2.0 * Z0 + ZpR) + dt * dt * sz * sx * ZpR;
  ZpR = Zsz + Rmz;
  double Sz0 = 16.0 * Z0 + 4.0 * ZpR + 2.0 * dt * (sx + sy) * (2.0 * Z0 + ZpR) + dt * dt * sx * sy * ZpR;
  double Sz1 = -32.0 * Z0 - 8.0 * ZpR + 2.0 * dt * dt * sx * sy * ZpR;
  double Sz2 = 16.0 * Z0 + 4.0 * ZpR - 2.0 * dt * (sx + sy) * (2.0 * Z0 + ZpR) + dt * dt * sx * sy * ZpR;
  double Tx0 = 8.0 + 4.0 * dt * sx, Tx1 = -16.0, Tx2 = 8.0 - 4.0 * dt * sx;
  double Ty0 = 8.0 + 4.0 * dt * sy, Ty1 = -16.0, Ty2 = 8.0 - 4.0 * dt * sy;
  double Tz0 = 8.0 + 4.0 * dt * sz, Tz1 = -16.0, Tz2 = 8.0 - 4.0 * dt * sz;
  double Uxy0 = 2.0 + dt * sz, Uxy1 = -2.0 + dt * sz, Wxy0 = Z0 * (2.0 + dt * sy), Wxy1 = Z0 * (-2.0 + dt * sy);
  double Uyx0 = 2.0 + dt * sz, Uyx1 = -2.0 + dt * sz, Wyx0 = Z0 * (2.0 + dt * sx), Wyx1 = Z0 * (-2.0 + dt * sx);
  double Uyz0 = 2.0 + dt * sx, Uyz1 = -2.0 + dt * sx, Wyz0 = Z0 * (2.0 + dt * sz), Wyz1 = Z0 * (-2.0 + dt * sz);
  double Uzy0 = 2.0 + dt * sx, Uzy1 = -2.0 + dt * sx, Wzy0 = Z0 * (2.0 + dt * sy), Wzy1 = Z0 * (-2.0 + dt * sy);
  double Uzx0 = 2.0 + dt * sy, Uzx1 = -2.0 + dt * sy, Wzx0 = Z0 * (2.0 + dt * sx), Wzx1 = Z0 * (-2.0 + dt * sx);
  double Uxz0 = 2.0 + dt * sy, Uxz1 = -2.0 + dt * sy, Wxz0 = Z0 * (2.0 + dt * sz), Wxz1 = Z0 * (-2.0 + dt * sz);
  double Px0 = 4.0 + 2.0 * dt * sx, Px1 = -8.0, Px2 = 4.0 - 2.0 * dt * sx;
  double Py0 = 4.0 + 2.0 * dt * sy, Py1 = -8.0, Py2 = 4.0 - 2.0 * dt * sy;
  double Pz0 = 4.0 + 2.0 * dt * sz, Pz1 = -8.0, Pz2 = 4.0 - 2.0 * dt * sz;
  double Qx0 = 4.0 * Zsx + 2.0 * dt * (sy + sz) * Zsx + 1.0 * (dt * dt * sy * sz * Zsx);
  double Qx1 = -8.0 * Zsx + 2.0 * (dt * dt * sy * sz * Zsx);
  double Qx2 = 4.0 * Zsx - 2.0 * dt * (sy + sz) * Zsx + 1.0 * (dt * dt * sy * sz * Zsx);
  double Qy0 = 4.0 * Zsy + 2.0 * dt * (sz + sx) * Zsy + 1.0 * (dt * dt * sz * sx * Zsy);
  double Qy1 = -8.0 * Zsy + 2.0 * (dt * dt * sz * sx * Zsy);
  double Qy2 = 4.0 * Zsy - 2.0 * dt * (sz + sx) * Zsy + 1.0 * (dt * dt * sz * sx * Zsy);
  double Qz0 = 4.0 * Zsz + 2.0 * dt * (sx + sy) * Zsz + 1.0 * (dt * dt * sx * sy * Zsz);
  double Qz1 = -8.0 * Zsz + 2.0 * (dt * dt * sx * sy * Zsz);
  double Qz2 = 4.0 * Zsz - 2.0 * dt * (sx + sy) * Zsz + 1.0 * (dt * dt * sx * sy * Zsz);
  /* STAGE 2 */
  double* S = &h->pml_state[(size_t)(PML_cell - 1) * PS_N];
  double* M = &h->pml_mat_state[(size_t)(h->pml_mat_pos[PML_cell - 1] - 1) * PM_N];
  double Ixt_m1 = S[PS_Ixt], Iyt_m1 = S[PS_Iyt], Izt_m1 = S[PS_Izt];
  double Vi_m1[13], Vi_m2[13];
  for (int i = 1; i <= 12; i++) Vi_m1[i] = S[PS_Vi + i - 1];
  double Vxt_m1 = S[PS_Vxt], Vyt_m1 = S[PS_Vyt], Vzt_m1 = S[PS_Vzt];
  double Vyxt_m1 = S[PS_Vyxt], Vzxt_m1 = S[PS_Vzxt], Vxyt_m1 = S[PS_Vxyt];
  double Vzyt_m1 = S[PS_Vzyt], Vyzt_m1 = S[PS_Vyzt], Vxzt_m1 = S[PS_Vxzt];
  double Ixt_m2 = M[PM_Ixt_m1], Iyt_m2 = M[PM_Iyt_m1], Izt_m2 = M[PM_Izt_m1];
  double Vcx_m2 = M[PM_Vcx_m1], Vcy_m2 = M[PM_Vcy_m1], Vcz_m2 = M[PM_Vcz_m1];
  double Vcx_m1 = M[PM_Vcx], Vcy_m1 = M[PM_Vcy], Vcz_m1 = M[PM_Vcz];
  for (int i = 1; i <= 12; i++) Vi_m2[i] = M[PM_Vi_m1 + i - 1];
  double Vxt_m2 = M[PM_Vxt_m1], Vyt_m2 = M[PM_Vyt_m1], Vzt_m2 = M[PM_Vzt_m1];
  double Vyxt_m2 = M[PM_Vyxt_m1], Vzxt_m2 = M[PM_Vzxt_m1], Vxyt_m2 = M[PM_Vxyt_m1];
  double Vzyt_m2 = M[PM_Vzyt_m1], Vyzt_m2 = M[PM_Vyzt_m1], Vxzt_m2 = M[PM_Vxzt_m1];
  (void)Vyxt_m2; (void)Vzxt_m2; (void)Vxyt_m2; (void)Vzyt_m2; (void)Vyzt_m2; (void)Vxzt_m2;
  double Vosx = M[PM_Vosxr], Vosy = M[PM_Vosyr], Vosz = M[PM_Voszr];
  double Vosx_m1 = M[PM_Vosx], Vosx_m2 = M[PM_Vosx_m1];
  double Vosy_m1 = M[PM_Vosy], Vosy_m2 = M[PM_Vosy_m1];
  double Vosz_m1 = M[PM_Vosz], Vosz_m2 = M[PM_Vosz_m1];
  double Vssx = -M[PM_Vssxr], Vssy = -M[PM_Vssyr], Vssz = -M[PM_Vsszr];
  double Vsx_m1 = M[PM_Vsx], Vsx_m2 = M[PM_Vsx_m1];
  double Vsy_m1 = M[PM_Vsy], Vsy_m2 = M[PM_Vsy_m1];
  double Vsz_m1 = M[PM_Vsz], Vsz_m2 = M[PM_Vsz_m1];
  /* STAGE 3 */
  v[Vynx] = v[Vynx] * exp_yi; v[Vypx] = v[Vypx] * exp_yi;
  v[Vznx] = v[Vznx] * exp_zi; v[Vzpx] = v[Vzpx] * exp_zi;
  v[Vxny] = v[Vxny] * exp_xi; v[Vxpy] = v[Vxpy] * exp_xi;
  v[Vzny] = v[Vzny] * exp_zi; v[Vzpy] = v[Vzpy] * exp_zi;
  v[Vxnz] = v[Vxnz] * exp_xi; v[Vxpz] = v[Vxpz] * exp_xi;
  v[Vynz] = v[Vynz] * exp_yi; v[Vypz] = v[Vypz] * exp_yi;
  /* STAGE 4 */
  double Vxt = (1.0 / Ax0) * (Bx0 * (v[Vznx] + v[Vzpx]) + Cx0 * (v[Vynx] + v[Vypx]) + Dx0 * (Vosx)
                              + Bx1 * (Vi_m1[Vznx] + Vi_m1[Vzpx]) + Cx1 * (Vi_m1[Vynx] + Vi_m1[Vypx]) + Dx1 * (Vosx_m1)
                              + Bx2 * (Vi_m2[Vznx] + Vi_m2[Vzpx]) + Cx2 * (Vi_m2[Vynx] + Vi_m2[Vypx]) + Dx2 * (Vosx_m2)
                              - Ax1 * Vxt_m1 - Ax2 * Vxt_m2);
  double Vyt = (1.0 / Ay0) * (By0 * (v[Vxny] + v[Vxpy]) + Cy0 * (v[Vzny] + v[Vzpy]) + Dy0 * (Vosy)
                              + By1 * (Vi_m1[Vxny] + Vi_m1[Vxpy]) + Cy1 * (Vi_m1[Vzny] + Vi_m1[Vzpy]) + Dy1 * (Vosy_m1)
                              + By2 * (Vi_m2[Vxny] + Vi_m2[Vxpy]) + Cy2 * (Vi_m2[Vzny] + Vi_m2[Vzpy]) + Dy2 * (Vosy_m2)
                              - Ay1 * Vyt_m1 - Ay2 * Vyt_m2);
  double Vzt = (1.0 / Az0) * (Bz0 * (v[Vynz] + v[Vypz]) + Cz0 * (v[Vxnz] + v[Vxpz]) + Dz0 * (Vosz)
                              + Bz1 * (Vi_m1[Vynz] + Vi_m1[Vypz]) + Cz1 * (Vi_m1[Vxnz] + Vi_m1[Vxpz]) + Dz1 * (Vosz_m1)
                              + Bz2 * (Vi_m2[Vynz] + Vi_m2[Vypz]) + Cz2 * (Vi_m2[Vxnz] + Vi_m2[Vxpz]) + Dz2 * (Vosz_m2)
                              - Az1 * Vzt_m1 - Az2 * Vzt_m2);
  /* STAGE 5 */
  double Vcx = v[Vypz] - v[Vzpy] - v[Vynz] + v[Vzny] - Vssx;
  double Vcy = v[Vzpx] - v[Vxpz] - v[Vznx] + v[Vxnz] - Vssy;
  double Vcz = v[Vxpy] - v[Vypx] - v[Vxny] + v[Vynx] - Vssz;
  double Ixt = (1.0 / Sx0) * (Tx0 * Vcx + Tx1 * Vcx_m1 + Tx2 * Vcx_m2 - Sx1 * Ixt_m1 - Sx2 * Ixt_m2);
  double Iyt = (1.0 / Sy0) * (Ty0 * Vcy + Ty1 * Vcy_m1 + Ty2 * Vcy_m2 - Sy1 * Iyt_m1 - Sy2 * Iyt_m2);
  double Izt = (1.0 / Sz0) * (Tz0 * Vcz + Tz1 * Vcz_m1 + Tz2 * Vcz_m2 - Sz1 * Izt_m1 - Sz2 * Izt_m2);
  double Vxyt = (1.0 / Uxy0) * (Wxy0 * Izt + Wxy1 * Izt_m1 - Uxy1 * Vxyt_m1);
  double Vyxt = (1.0 / Uyx0) * (Wyx0 * Izt + Wyx1 * Izt_m1 - Uyx1 * Vyxt_m1);
  double Vyzt = (1.0 / Uyz0) * (Wyz0 * Ixt + Wyz1 * Ixt_m1 - Uyz1 * Vyzt_m1);
  double Vzyt = (1.0 / Uzy0) * (Wzy0 * Ixt + Wzy1 * Ixt_m1 - Uzy1 * Vzyt_m1);
  double Vxzt = (1.0 / Uxz0) * (Wxz0 * Iyt + Wxz1 * Iyt_m1 - Uxz1 * Vxzt_m1);
  double Vzxt = (1.0 / Uzx0) * (Wzx0 * Iyt + Wzx1 * Iyt_m1 - Uzx1 * Vzxt_m1);
  /* STAGE 9: the previous step's values move to the _m1 slots */
  for (int i = 1; i <= 12; i++) { M[PM_Vi_m1 + i - 1] = S[PS_Vi + i - 1]; S[PS_Vi + i - 1] = v[i]; }
  M[PM_Vxt_m1] = S[PS_Vxt]; M[PM_Vyt_m1] = S[PS_Vyt]; M[PM_Vzt_m1] = S[PS_Vzt];
  S[PS_Vxt] = Vxt; S[PS_Vyt] = Vyt; S[PS_Vzt] = Vzt;
  M[PM_Ixt_m1] = S[PS_Ixt]; M[PM_Iyt_m1] = S[PS_Iyt]; M[PM_Izt_m1] = S[PS_Izt];
  S[PS_Ixt] = Ixt; S[PS_Iyt] = Iyt; S[PS_Izt] = Izt;
  M[PM_Vcx_m1] = M[PM_Vcx]; M[PM_Vcy_m1] = M[PM_Vcy]; M[PM_Vcz_m1] = M[PM_Vcz];
  M[PM_Vcx] = Vcx; M[PM_Vcy] = Vcy; M[PM_Vcz] = Vcz;
  M[PM_Vxyt_m1] = S[PS_Vxyt]; M[PM_Vyxt_m1] = S[PS_Vyxt];
  S[PS_Vxyt] = Vxyt; S[PS_Vyxt] = Vyxt;
  M[PM_Vzxt_m1] = S[PS_Vzxt]; M[PM_Vxzt_m1] = S[PS_Vxzt];
  S[PS_Vzxt] = Vzxt; S[PS_Vxzt] = Vxzt;
  M[PM_Vzyt_m1] = S[PS_Vzyt]; M[PM_Vyzt_m1] = S[PS_Vyzt];
  S[PS_Vzyt] = Vzyt; S[PS_Vyzt] = Vyzt;
  /* STAGE 11 */
  double V_max, V_min;
  V_max = v[Vypx]; V_min = v[Vynx];
  v[Vynx] = Vxt - Vyxt - V_max;
  v[Vypx] = Vxt + Vyxt - V_min;
  V_max = v[Vzpx]; V_min = v[Vznx];
  v[Vznx] = Vxt + Vzxt - V_max;
  v[Vzpx] = Vxt - Vzxt - V_min;
  V_max = v[Vzpy]; V_min = v[Vzny];
  v[Vzny] = Vyt - Vzyt - V_max;
  v[Vzpy] = Vyt + Vzyt - V_min;
  V_max = v[Vxpy]; V_min = v[Vxny];
  v[Vxny] = Vyt + Vxyt - V_max;
  v[Vxpy] = Vyt - Vxyt - V_min;
  V_max = v[Vxpz]; V_min = v[Vxnz];
  v[Vxnz] = Vzt - Vxzt - V_max;
  v[Vxpz] = Vzt + Vxzt - V_min;
  V_max = v[Vypz]; V_min = v[Vynz];
  v[Vynz] = Vzt + Vyzt - V_max;
  v[Vypz] = Vzt - Vyzt - V_min;
  /* capacitive stub */
  M[PM_Vosxr] = Vxt - Vosx; M[PM_Vosyr] = Vyt - Vosy; M[PM_Voszr] = Vzt - Vosz;
  M[PM_Vosx_m1] = M[PM_Vosx]; M[PM_Vosx] = Vosx;
  M[PM_Vosy_m1] = M[PM_Vosy]; M[PM_Vosy] = Vosy;
  M[PM_Vosz_m1] = M[PM_Vosz]; M[PM_Vosz] = Vosz;
  /* inductive stub */
  double Vsx = (1.0 / Px0) * (Qx0 * Ixt + Qx1 * Ixt_m1 + Qx2 * Ixt_m2 - Px1 * Vsx_m1 - Px2 * Vsx_m2);
  double Vsy = (1.0 / Py0) * (Qy0 * Iyt + Qy1 * Iyt_m1 + Qy2 * Iyt_m2 - Py1 * Vsy_m1 - Py2 * Vsy_m2);
  double Vsz = (1.0 / Pz0) * (Qz0 * Izt + Qz1 * Izt_m1 + Qz2 * Izt_m2 - Pz1 * Vsz_m1 - Pz2 * Vsz_m2);
  M[PM_Vssxr] = (Vsx + Vssx); M[PM_Vssyr] = (Vsy + Vssy); M[PM_Vsszr] = (Vsz + Vssz);
  M[PM_Vsx_m1] = M[PM_Vsx]; M[PM_Vsx] = Vsx;
  M[PM_Vsy_m1] = M[PM_Vsy]; M[PM_Vsy] = Vsy;
  M[PM_Vsz_m1] = M[PM_Vsz]; M[PM_Vsz] = Vsz;
  /* STAGE 12 */
  if (output_number != 0) {
    *output_cell_number += 1;
    double* fo = &h->cell_out[(size_t)(*output_cell_number - 1) * 6];
    fo[0] = -Vxt / dl; fo[1] = -Vyt / dl; fo[2] = -Vzt / dl;
    fo[3] = Ixt / dl; fo[4] = Iyt / dl; fo[5] = Izt / dl;
  }
  /* STAGE 13 */
  v[Vynx] = v[Vynx] * exp_yr; v[Vypx] = v[Vypx] * exp_yr;
  v[Vznx] = v[Vznx] * exp_zr; v[Vzpx] = v[Vzpx] * exp_zr;
  v[Vxny] = v[Vxny] * exp_xr; v[Vxpy] = v[Vxpy] * exp_xr;
  v[Vzny] = v[Vzny] * exp_zr; v[Vzpy] = v[Vzpy] * exp_zr;
  v[Vxnz] = v[Vxnz] * exp_xr; v[Vxpz] = v[Vxpz] * exp_xr;
  v[Vynz] = v[Vynz] * exp_yr; v[Vypz] = v[Vypz] * exp_yr;
}

/* frequency_output_volumes.F90:309-386 (kind 1, the cell outputs of the
   step, time (t-1) dt) and frequency_output_surfaces.F90:297-421 (kind 2,
   the face outputs of the step's connect, time (t-1) dt + dt/2, J = n x H
   for components 7..9): value * exp(j y), y = -2 pi f time, added to each
   item of every output due at this timestep (get_output_flag,
   output.F90:564-). */
static void accumulate_frequency_outputs(oracle_handle* h, int kind, const double* src, int timestep,
                                         double time) {
  for (int i = 0; i < h->n_fo_items; i++) {
    const tlm_frequency_output* o = &h->fo[h->fo_out[i]];
    if (o->kind != kind || timestep < o->first_timestep || timestep > o->last_timestep ||
        (timestep - o->first_timestep) % o->timestep_interval != 0)
      continue;
    double value;
    if (kind == 1) {
      value = src[(size_t)(h->fo_row[i] - 1) * 6 + (o->field_component - 1)];
    } else {
      const double* f = &src[((size_t)(h->fo_row[i] - 1) * 2 + (h->fo_side[i] - 1)) * 6];
      int nrm = h->fo_norm[i];
      double normx = nrm == 1 ? 1.0 : nrm == -1 ? -1.0 : 0.0;
      double normy = nrm == 2 ? 1.0 : nrm == -2 ? -1.0 : 0.0;
      double normz = nrm == 3 ? 1.0 : nrm == -3 ? -1.0 : 0.0;
      if (o->field_component <= 6) value = f[o->field_component - 1];
      else if (o->field_component == 7) value = normy * f[5] - normz * f[4];
      else if (o->field_component == 8) value = normz * f[3] - normx * f[5];
      else value = normx * f[4] - normy * f[3];
    }
    double y = -(2.0 * PI * o->frequency * time);
    h->fo_acc[2 * i] += value * cos(y);
    h->fo_acc[2 * i + 1] += value * sin(y);
  }
}

/* face_output_surfaces (output_surfaces.F90:601-720) for kind 2 and
   cell_output_volumes (output_volumes.F90:496-545) for kind 1: every output
   due at this timestep (get_output_flag) records one value per item. */
static void record_field_outputs(oracle_handle* h, int kind, const double* src, int timestep) {
  size_t k = 0;
  for (int o = 0; o < h->n_fd; o++) {
    const tlm_field_output* f = &h->fd[o];
    size_t k0 = k;
    k += (size_t)f->n_items;
    if (f->kind != kind || timestep < f->first_timestep || timestep > f->last_timestep ||
        (timestep - f->first_timestep) % f->timestep_interval != 0)
      continue;
    if (h->fd_nframes[o] == h->fd_capframes[o]) {
      h->fd_capframes[o] = h->fd_capframes[o] ? 2 * h->fd_capframes[o] : 16;
      h->fd_frames[o] = (double*)realloc(h->fd_frames[o],
                                         (size_t)h->fd_capframes[o] * (f->n_items ? f->n_items : 1) * sizeof(double));
    }
    if (h->fd_nframes[o] == 0) h->fd_first[o] = timestep;
    double* out = h->fd_frames[o] + (size_t)h->fd_nframes[o] * f->n_items;
    int c = f->field_component;
    for (int i = 0; i < f->n_items; i++) {
      size_t q = k0 + (size_t)i;
      double value;
      if (kind == 1) {
        value = src[(size_t)(h->fd_row[q] - 1) * 6 + (c - 1)];
      } else {
        int nrm = h->fd_norm[q];
        double normx = nrm == 1 ? 1.0 : nrm == -1 ? -1.0 : 0.0;
        double normy = nrm == 2 ? 1.0 : nrm == -2 ? -1.0 : 0.0;
        double normz = nrm == 3 ? 1.0 : nrm == -3 ? -1.0 : 0.0;
        double field[6], field2[6];
        const double* blk = &src[(size_t)(h->fd_row[q] - 1) * 12];
        if (!f->both_sides) {
          memcpy(field, blk + (size_t)(h->fd_side[q] - 1) * 6, sizeof(field));
          for (int j = 0; j < 6; j++) field2[j] = 0.0;
        } else {
          memcpy(field, blk, sizeof(field));
          memcpy(field2, blk + 6, sizeof(field2));
        }
        double Jsx = (normy * field[5] - normz * field[4]) - (normy * field2[5] - normz * field2[4]);
        double Jsy = (normz * field[3] - normx * field[5]) - (normz * field2[3] - normx * field2[5]);
        double Jsz = (normx * field[4] - normy * field[3]) - (normx * field2[4] - normy * field2[3]);
        if (c <= 6) value = field[c - 1];
        else if (c == 7) value = Jsx;
        else if (c == 8) value = Jsy;
        else if (c == 9) value = Jsz;
        else if (c == 13) value = sqrt(field[0] * field[0] + field[1] * field[1] + field[2] * field[2]);
        else if (c == 14) value = sqrt(field[3] * field[3] + field[4] * field[4] + field[5] * field[5]);
        else value = sqrt(Jsx * Jsx + Jsy * Jsy + Jsz * Jsz);
      }
      out[i] = value;
    }
    h->fd_nframes[o]++;
  }
}

/* face_output_far_field (far_field_surfaces.F90:252-351): every connect
   adds ejwt (n x H) to J and -ejwt (n x E) to M of each face, ejwt =
   exp(-j 2 pi f time) dt, time = (t-1) dt + dt/2, with the field of the
   face's side (1 for a min face, 2 for a max face) and its outward normal. */
static void far_field_currents(oracle_handle* h, const double* src, double time) {
  size_t k = 0;
  for (int s = 0; s < h->n_ff; s++) {
    const tlm_far_field_surface* f = &h->ff[s];
    double y = -(2.0 * PI * f->frequency * time);
    double er = cos(y) * h->g.dt, ei = sin(y) * h->g.dt;
    for (int i = 0; i < f->n_faces; i++, k++) {
      const double* fld = &src[((size_t)(h->ff_row[k] - 1) * 2 + (h->ff_side[k] - 1)) * 6];
      int nrm = h->ff_norm[k];
      double normx = nrm == 1 ? 1.0 : nrm == -1 ? -1.0 : 0.0;
      double normy = nrm == 2 ? 1.0 : nrm == -2 ? -1.0 : 0.0;
      double normz = nrm == 3 ? 1.0 : nrm == -3 ? -1.0 : 0.0;
      double jv[3] = {normy * fld[5] - normz * fld[4], normz * fld[3] - normx * fld[5], normx * fld[4] - normy * fld[3]};
      double mv[3] = {normy * fld[2] - normz * fld[1], normz * fld[0] - normx * fld[2], normx * fld[1] - normy * fld[0]};
      for (int c = 0; c < 3; c++) {
        double* J = &h->ff_J[(k * 3 + c) * 2];
        double* M = &h->ff_M[(k * 3 + c) * 2];
        J[0] = J[0] + er * jv[c];
        J[1] = J[1] + ei * jv[c];
        M[0] = M[0] - er * mv[c];
        M[1] = M[1] - ei * mv[c];
      }
    }
  }
}

/* face_output_rcs (RCS_surfaces.F90:386-559): each face's far field in the
   monostatic direction, added to the far-field point of its retarded
   distance c0 time - K.r, faces in the caller's order. */
static void rcs_currents(oracle_handle* h, const double* src, double time) {
  const double dl = h->g.dl, dA = dl * dl;
  double theta = h->rcs.theta, phi = h->rcs.phi;
  double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
  double K[3] = {1.0 * sin(theta) * cos(phi), 1.0 * sin(theta) * sin(phi), 1.0 * cos(theta)};
  for (int i = 0; i < h->rcs.n_faces; i++) {
    const double* fld = &src[((size_t)(h->rcs_row[i] - 1) * 2 + (h->rcs_side[i] - 1)) * 6];
    int nrm = h->rcs_norm[i];
    double normx = nrm == 1 ? 1.0 : nrm == -1 ? -1.0 : 0.0;
    double normy = nrm == 2 ? 1.0 : nrm == -2 ? -1.0 : 0.0;
    double normz = nrm == 3 ? 1.0 : nrm == -3 ? -1.0 : 0.0;
    double Jsx = (normy * fld[5] - normz * fld[4]);
    double Jsy = (normz * fld[3] - normx * fld[5]);
    double Jsz = (normx * fld[4] - normy * fld[3]);
    double Msx = -(normy * fld[2] - normz * fld[1]);
    double Msy = -(normz * fld[0] - normx * fld[2]);
    double Msz = -(normx * fld[1] - normy * fld[0]);
    const double* r = &h->rcs_centre[3 * (size_t)i];
    double dmrcs = K[0] * r[0] + K[1] * r[1] + K[2] * r[2];
    dmrcs = -dmrcs + time * C0;
    int point = (int)((double)h->rcs.points_per_cell * dmrcs / dl);
    point = point - h->rcs.point_offset;
    if (point >= 1 && point <= h->rcs.n_points) {
      double Ntheta = (Jsx * ct * cp + Jsy * ct * sp - Jsz * st) * dA / (4.0 * PI);
      double Nphi = (-Jsx * sp + Jsy * cp) * dA / (4.0 * PI);
      double Ltheta = (Msx * ct * cp + Msy * ct * sp - Msz * st) * dA / (4.0 * PI);
      double Lphi = (-Msx * sp + Msy * cp) * dA / (4.0 * PI);
      double* E = &h->rcs_E[4 * (size_t)(point - 1)];
      E[0] = E[0] + -(Lphi + Z0 * Ntheta);
      E[1] = E[1] + (Ltheta - Z0 * Nphi);
      E[2] = E[2] + (Nphi - Ltheta / Z0);
      E[3] = E[3] + -(Ntheta + Lphi / Z0);
    }
  }
}

static void record_outputs(oracle_handle* h, int timestep) {
  /* cell_output(): copy this step's cell_output_field into the ring and
   * accumulate the probe DFT at time=(timestep-1)*dt (run_TLM.F90:240). */
  int rows = h->n_out_rows;
  if (!h->ring && rows) h->ring = (double*)malloc((size_t)h->cap * rows * 6 * sizeof(double));
  if (rows) {
    if (h->ring_count == 0) h->ring_first = timestep;
    if (h->ring_count < h->cap) {
      memcpy(h->ring + (size_t)h->ring_count * rows * 6, h->cell_out, (size_t)rows * 6 * sizeof(double));
      h->ring_count++;
    }
  }
  if (h->n_freq && h->n_points) {
    double dt = h->g.dt;
    double time = (double)(timestep - 1) * dt;
    for (int p = 0; p < h->n_points; p++) {
      double v = h->cell_out[(size_t)(h->pt_row[p] - 1) * 6 + (h->pt_comp[p] - 1)];
      for (int k = 0; k < h->n_freq; k++) {
        double ph = 2.0 * PI * h->freq[k] * time;
        double* acc = &h->dft[((size_t)p * h->n_freq + k) * 2];
        acc[0] += v * cos(ph) * dt;
        acc[1] -= v * sin(ph) * dt;
      }
    }
  }
  if (h->n_fo_items && rows) accumulate_frequency_outputs(h, 1, h->cell_out, timestep, (double)(timestep - 1) * h->g.dt);
  if (h->n_fd && rows) record_field_outputs(h, 1, h->cell_out, timestep);
}

/* Scan-order counters of scatter() at the start of each z-plane: output
 * cells (generic, PML and PML-material cells all number their rows) and
 * excitation cells (counted by the Thevenin path only, scatter.F90:431). */
static int cell_prefixes(oracle_handle* h) {
  if (h->pl_out) return TLM_OK;
  int nx = h->nx, ny = h->ny, nz = h->nz;
  h->pl_out = (int32_t*)calloc((size_t)nz + 1, sizeof(int32_t));
  h->pl_exc = (int32_t*)calloc((size_t)nz + 1, sizeof(int32_t));
  if (!h->pl_out || !h->pl_exc) return fail(h, TLM_ERR_NOMEM, "plane counters");
  int no = 0, ne = 0;
  size_t c = 0;
  for (int cz = 1; cz <= nz; cz++) {
    h->pl_out[cz - 1] = no;
    h->pl_exc[cz - 1] = ne;
    for (int k = 0; k < nx * ny; k++, c++) {
      int code = h->cell_code[c];
      if (!code || h->c_cable[code - 1]) continue;
      no += h->c_out[code - 1] != 0;
      ne += h->c_pml[code - 1] == 0 && h->c_exc[code - 1] != 0;
    }
  }
  h->pl_out[nz] = no;
  h->pl_exc[nz] = ne;
  return TLM_OK;
}

/* scatter() over the cells of z-plane cz (scatter.F90:216-585 restricted to
 * one plane); the planes are independent, a plane's counters start from
 * cell_prefixes. */
static int scatter_plane(oracle_handle* h, int cz, int32_t timestep) {
  const double dl = h->g.dl;
  const double Z04 = Z0 * 4.0, Y04 = Y0 * 4.0;
  int nx = h->nx, ny = h->ny;
  int cell_number = (cz - 1) * nx * ny;
  int output_cell_number = h->pl_out[cz - 1], excitation_cell_number = h->pl_exc[cz - 1];
  /* TLM_set_small_to_zero.F90:34-68 at the top of a step whose number is a
   * multiple of n_small_to_zero (run_TLM.F90:236-238); excitation() does
   * not touch V, and the flush and the scatter are cell-local, so flushing
   * this plane just before its scatter is the same */
  if (h->stz_n > 0 && timestep % h->stz_n == 0)
    for (int cy = 1; cy <= ny; cy++)
      for (int cx = 1; cx <= nx; cx++) {
        double* v = cellV(h, cx, cy, cz);
        for (int i = 1; i <= 12; i++)
          if (fabs(v[i]) < h->stz_value) v[i] = 0.0;
      }
  {
    for (int cy = 1; cy <= ny; cy++)
      for (int cx = 1; cx <= nx; cx++) {
        cell_number++;
        double* v = cellV(h, cx, cy, cz);
        int code = h->cell_code[cell_number - 1];
        if (code == 0) {
          double Vx = (v[Vx_ymin] + v[Vx_ymax] + v[Vx_zmin] + v[Vx_zmax]) / 2.0;
          double Vy = (v[Vy_xmin] + v[Vy_xmax] + v[Vy_zmin] + v[Vy_zmax]) / 2.0;
          double Vz = (v[Vz_xmin] + v[Vz_xmax] + v[Vz_ymin] + v[Vz_ymax]) / 2.0;
          double Ix = (v[Vz_ymax] - v[Vy_zmax] - v[Vz_ymin] + v[Vy_zmin]) / (2.0 * Z0);
          double Iy = (v[Vx_zmax] - v[Vz_xmax] - v[Vx_zmin] + v[Vz_xmin]) / (2.0 * Z0);
          double Iz = (v[Vy_xmax] - v[Vx_ymax] - v[Vy_xmin] + v[Vx_ymin]) / (2.0 * Z0);
          scatter_ports(v, Vx, Vy, Vz, Ix, Iy, Iz);
          continue;
        }
        int special = code;
        /* host-path (cable) cell: left incident for the host's own scatter
           (scatter.F90:290,453-462 run the cable junction inside this update) */
        if (h->c_cable[special - 1]) continue;
        int PML_cell = h->c_pml[special - 1];
        int material_number = h->c_mat[special - 1];
        int material_type = material_number ? h->vmat[material_number - 1].type : 0;
        int output_number = h->c_out[special - 1];
        if (PML_cell != 0) {
          if (material_type == TLM_VMAT_PEC || material_type == TLM_VMAT_PMC)
            return fail(h, TLM_ERR_ARG, "PEC/PMC volume material in PML");
          if (material_type == TLM_VMAT_DISPERSIVE) {
            /* PML material cells were numbered before the planes ran */
            scatter_pml_material_cell(h, v, PML_cell, material_number, output_number, &output_cell_number);
            continue;
          }
          scatter_pml_cell(h, v, PML_cell, output_number, &output_cell_number);
          continue;
        }
        int excitation_number = h->c_exc[special - 1];
        const tlm_volume_material* vm = material_number ? &h->vmat[material_number - 1] : NULL;
        double VEx = (v[Vx_ymin] + v[Vx_ymax] + v[Vx_zmin] + v[Vx_zmax]) / 2.0;
        double VEy = (v[Vy_xmin] + v[Vy_xmax] + v[Vy_zmin] + v[Vy_zmax]) / 2.0;
        double VEz = (v[Vz_xmin] + v[Vz_xmax] + v[Vz_ymin] + v[Vz_ymax]) / 2.0;
        double Yx = Y04, Yy = Y04, Yz = Y04;
        double VHx = 2.0 * (v[Vz_ymax] - v[Vy_zmax] - v[Vz_ymin] + v[Vy_zmin]);
        double VHy = 2.0 * (v[Vx_zmax] - v[Vz_xmax] - v[Vx_zmin] + v[Vz_xmin]);
        double VHz = 2.0 * (v[Vy_xmax] - v[Vx_ymax] - v[Vy_xmin] + v[Vx_ymin]);
        double Zx = Z04, Zy = Z04, Zz = Z04;
        double Vox = 0, Voy = 0, Voz = 0, Ysx = 0, Ysy = 0, Ysz = 0;
        int filter_number = 0;
        if (material_type == TLM_VMAT_PEC) {
          VEx = 0.0; VEy = 0.0; VEz = 0.0;
        } else if (material_type == TLM_VMAT_PMC) {
          VHx = 0.0; VHy = 0.0; VHz = 0.0;
        } else if (material_type == TLM_VMAT_DISPERSIVE) {
          filter_number = h->c_filt[special - 1];
          if (vm->eps_filter_exists) {
            Ysx = vm->Ys_eps_f; Ysy = vm->Ys_eps_f; Ysz = vm->Ys_eps_f;
            int o = vm->eps_order;
            double* w;
            w = &h->eps_w[(size_t)(filter_number) * NW];
            timeshift_Zfilter(w, o);
            Vox = evaluate_Zfilter(vm->eps_a, vm->eps_b, w, o, 0.0);
            w = &h->eps_w[(size_t)(filter_number + 1) * NW];
            timeshift_Zfilter(w, o);
            Voy = evaluate_Zfilter(vm->eps_a, vm->eps_b, w, o, 0.0);
            w = &h->eps_w[(size_t)(filter_number + 2) * NW];
            timeshift_Zfilter(w, o);
            Voz = evaluate_Zfilter(vm->eps_a, vm->eps_b, w, o, 0.0);
            add_shunt(&VEx, &Yx, Vox, Ysx);
            add_shunt(&VEy, &Yy, Voy, Ysy);
            add_shunt(&VEz, &Yz, Voz, Ysz);
          }
          double Gex = vm->Ge, Gey = vm->Ge, Gez = vm->Ge;
          add_shunt(&VEx, &Yx, 0.0, Gex);
          add_shunt(&VEy, &Yy, 0.0, Gey);
          add_shunt(&VEz, &Yz, 0.0, Gez);
          if (vm->mu_filter_exists) {
            double Zsx = vm->Zs_mu_f, Zsy = vm->Zs_mu_f, Zsz = vm->Zs_mu_f;
            int o = vm->mu_order;
            double* w;
            double Vsx, Vsy, Vsz;
            w = &h->mu_w[(size_t)(filter_number) * NW];
            timeshift_Zfilter(w, o);
            Vsx = evaluate_Zfilter(vm->mu_a, vm->mu_b, w, o, 0.0);
            w = &h->mu_w[(size_t)(filter_number + 1) * NW];
            timeshift_Zfilter(w, o);
            Vsy = evaluate_Zfilter(vm->mu_a, vm->mu_b, w, o, 0.0);
            w = &h->mu_w[(size_t)(filter_number + 2) * NW];
            timeshift_Zfilter(w, o);
            Vsz = evaluate_Zfilter(vm->mu_a, vm->mu_b, w, o, 0.0);
            add_series(&VHx, &Zx, Vsx, Zsx);
            add_series(&VHy, &Zy, Vsy, Zsy);
            add_series(&VHz, &Zz, Vsz, Zsz);
          }
          double Rmx = vm->Rm, Rmy = vm->Rm, Rmz = vm->Rm;
          add_series(&VHx, &Zx, 0.0, Rmx);
          add_series(&VHy, &Zy, 0.0, Rmy);
          add_series(&VHz, &Zz, 0.0, Rmz);
        }
        double Vx = VEx, Vy = VEy, Vz = VEz;
        double Ix = VHx / Zx, Iy = VHy / Zy, Iz = VHz / Zz;
        if (excitation_number != 0) {
          excitation_cell_number++;
          const double* field = &h->cell_exc_field[(size_t)(excitation_cell_number - 1) * 9];
          const int32_t* hsf = &h->cell_exc_type[(size_t)(excitation_cell_number - 1) * 6];
          double Isx = field[Jx - 1] * dl * dl;
          double Isy = field[Jy - 1] * dl * dl;
          double Isz = field[Jz - 1] * dl * dl;
          Vx = (double)hsf[0] * Vx - field[Ex - 1] * dl / 2.0 + Isx / Yx;
          Vy = (double)hsf[1] * Vy - field[Ey - 1] * dl / 2.0 + Isy / Yy;
          Vz = (double)hsf[2] * Vz - field[Ez - 1] * dl / 2.0 + Isz / Yz;
          Ix = (double)hsf[3] * Ix + field[Hx - 1] * dl / 2.0;
          Iy = (double)hsf[4] * Iy + field[Hy - 1] * dl / 2.0;
          Iz = (double)hsf[5] * Iz + field[Hz - 1] * dl / 2.0;
        }
        if (output_number != 0) {
          output_cell_number++;
          double* fo = &h->cell_out[(size_t)(output_cell_number - 1) * 6];
          fo[0] = -Vx / dl; fo[1] = -Vy / dl; fo[2] = -Vz / dl;
          fo[3] = Ix / dl; fo[4] = Iy / dl; fo[5] = Iz / dl;
        }
        if (material_type == TLM_VMAT_DISPERSIVE) {
          if (vm->eps_filter_exists) {
            double Iox = (Vx - Vox) * Ysx;
            double Ioy = (Vy - Voy) * Ysy;
            double Ioz = (Vz - Voz) * Ysz;
            int o = vm->eps_order;
            evaluate_Zfilter(vm->eps_a, vm->eps_b, &h->eps_w[(size_t)(filter_number) * NW], o, Iox);
            evaluate_Zfilter(vm->eps_a, vm->eps_b, &h->eps_w[(size_t)(filter_number + 1) * NW], o, Ioy);
            evaluate_Zfilter(vm->eps_a, vm->eps_b, &h->eps_w[(size_t)(filter_number + 2) * NW], o, Ioz);
          }
          if (vm->mu_filter_exists) {
            int o = vm->mu_order;
            evaluate_Zfilter(vm->mu_a, vm->mu_b, &h->mu_w[(size_t)(filter_number) * NW], o, -Ix);
            evaluate_Zfilter(vm->mu_a, vm->mu_b, &h->mu_w[(size_t)(filter_number + 1) * NW], o, -Iy);
            evaluate_Zfilter(vm->mu_a, vm->mu_b, &h->mu_w[(size_t)(filter_number + 2) * NW], o, -Iz);
          }
        }
        scatter_ports(v, Vx, Vy, Vz, Ix, Iy, Iz);
      }
  }
  return TLM_OK;
}

int oracle_scatter(oracle_handle* h, int32_t timestep) {
  if (!h->cell_code) return fail(h, TLM_ERR_STATE, "cell codes not set");
  int rc;
  if ((rc = cell_prefixes(h))) return rc;
  /* allocate_PML_material_data before the planes run (it numbers the PML
   * material cells in scan order) */
  if (h->n_pml_cells > 0 && (rc = setup_pml_materials(h))) return rc;
  int err = TLM_OK;
#pragma omp parallel for schedule(static) num_threads(h->threads)
  for (int cz = 1; cz <= h->nz; cz++) {
    int r = scatter_plane(h, cz, timestep);
    if (r) {
#pragma omp atomic write
      err = r;
    }
  }
  if (err) return err;
  record_outputs(h, timestep);
  return TLM_OK;
}

/* ---- connect ------------------------------------------------------------- */

/* One special face, connect_normal_to_{x,y,z}.F90 lines 12-361.
 * axis 0/1/2 = x/y/z normal.  pa/pb are the two polarisation ports on the
 * min side (cell c) and qa/qb the matching ports on the max side (cell c-1):
 * x: (Vy_xmin,Vz_xmin)/(Vy_xmax,Vz_xmax); y: (Vx_ymin,Vz_ymin)/(Vx_ymax,Vz_ymax);
 * z: (Vx_zmin,Vy_zmin)/(Vx_zmax,Vy_zmax). */
static int connect_special_face(oracle_handle* h, int axis, double* vmin_cell, double* vmax_cell,
                                int code, int* excitation_face_number, int* output_face_number) {
  static const int pa[3] = {Vy_xmin, Vx_ymin, Vx_zmin}, pb[3] = {Vz_xmin, Vz_ymin, Vy_zmin};
  static const int qa[3] = {Vy_xmax, Vx_ymax, Vx_zmax}, qb[3] = {Vz_xmax, Vz_ymax, Vy_zmax};
  const double dl = h->g.dl;
  int m = h->f_mat[code - 1];
  int material_number = m < 0 ? -m : m;
  int material_type = material_number ? h->smat[material_number - 1] : 0;
  int output_number = h->f_out[code - 1];
  int excitation_number = h->f_exc[code - 1];
  if (excitation_number < 0) excitation_number = -excitation_number;
  double* A = &vmin_cell[pa[axis]]; /* min-side first polarisation  */
  double* B = &vmin_cell[pb[axis]]; /* min-side second polarisation */
  double* C = &vmax_cell[qa[axis]]; /* max-side first polarisation  */
  double* D = &vmax_cell[qb[axis]]; /* max-side second polarisation */
  double a_in = *A, b_in = *B, c_in = *C, d_in = *D; /* pre-connect values */
  double a_min, a_max, b_min, b_max;
  if (material_type == 0 || material_type == TLM_SMAT_FREE_SPACE) {
    a_min = *A + *C; a_max = a_min;
    b_min = *B + *D; b_max = b_min;
  } else if (material_type == TLM_SMAT_PEC) {
    a_min = 0.0; a_max = 0.0; b_min = 0.0; b_max = 0.0;
  } else if (material_type == TLM_SMAT_PMC) {
    a_min = 2.0 * *A; a_max = 2.0 * *C;
    b_min = 2.0 * *B; b_max = 2.0 * *D;
  } else if (host_smat(material_type)) {
    /* the face voltages the host's material model computed from this
     * face's scattered pulses (tlm_scatter_host_faces) */
    const double* T = &h->hf_tot[(size_t)(h->hf_of_code[code - 1] - 1) * 4];
    a_min = T[0]; a_max = T[1];
    b_min = T[2]; b_max = T[3];
  } else {
    return fail(h, TLM_ERR_UNSUPPORTED, "surface material type stays on the host path");
  }
  if (excitation_number != 0) {
    *excitation_face_number += 1;
    const double* fmin = &h->face_exc_field[((size_t)(*excitation_face_number - 1) * 2 + 0) * 12];
    const double* fmax = &h->face_exc_field[((size_t)(*excitation_face_number - 1) * 2 + 1) * 12];
    /* hard_source_factor_max takes side 1's types (connect_normal_to_x.F90:294) */
    const int32_t* hmin = &h->face_exc_type[((size_t)(*excitation_face_number - 1) * 2 + 0) * 6];
    const int32_t* hmax = hmin;
    const double* Jsmin = fmin + 6; const double* Jsmax = fmax + 6;
    const double* Msmin = fmin + 9; const double* Msmax = fmax + 9;
    if (axis == 0) { /* connect_normal_to_x.F90:300-304: (a,b)=(y,z) */
      a_min = a_min * (double)hmin[Ey - 1] - fmin[Ey - 1] * dl - fmin[Hz - 1] * Z0 * dl + Jsmin[1] * dl * Z0 / 2.0 + Msmin[2] * dl / 2.0;
      a_max = a_max * (double)hmax[Ey - 1] - fmax[Ey - 1] * dl + fmax[Hz - 1] * Z0 * dl + Jsmax[1] * dl * Z0 / 2.0 - Msmax[2] * dl / 2.0;
      b_min = b_min * (double)hmin[Ez - 1] - fmin[Ez - 1] * dl + fmin[Hy - 1] * Z0 * dl + Jsmin[2] * dl * Z0 / 2.0 - Msmin[1] * dl / 2.0;
      b_max = b_max * (double)hmax[Ez - 1] - fmax[Ez - 1] * dl - fmax[Hy - 1] * Z0 * dl + Jsmax[2] * dl * Z0 / 2.0 + Msmax[1] * dl / 2.0;
    } else if (axis == 1) { /* connect_normal_to_y.F90:306-310: (a,b)=(x,z) */
      a_min = a_min * (double)hmin[Ex - 1] - fmin[Ex - 1] * dl + fmin[Hz - 1] * Z0 * dl + Jsmin[0] * dl * Z0 / 2.0 - Msmin[2] * dl / 2.0;
      a_max = a_max * (double)hmax[Ex - 1] - fmax[Ex - 1] * dl - fmax[Hz - 1] * Z0 * dl + Jsmax[0] * dl * Z0 / 2.0 + Msmax[2] * dl / 2.0;
      b_min = b_min * (double)hmin[Ez - 1] - fmin[Ez - 1] * dl - fmin[Hx - 1] * Z0 * dl + Jsmin[2] * dl * Z0 / 2.0 + Msmin[0] * dl / 2.0;
      b_max = b_max * (double)hmax[Ez - 1] - fmax[Ez - 1] * dl + fmax[Hx - 1] * Z0 * dl + Jsmax[2] * dl * Z0 / 2.0 - Msmax[0] * dl / 2.0;
    } else { /* connect_normal_to_z.F90:303-307: (a,b)=(x,y) */
      a_min = a_min * (double)hmin[Ex - 1] - fmin[Ex - 1] * dl - fmin[Hy - 1] * Z0 * dl + Jsmin[0] * dl * Z0 / 2.0 + Msmin[1] * dl / 2.0;
      a_max = a_max * (double)hmax[Ex - 1] - fmax[Ex - 1] * dl + fmax[Hy - 1] * Z0 * dl + Jsmax[0] * dl * Z0 / 2.0 - Msmax[1] * dl / 2.0;
      b_min = b_min * (double)hmin[Ey - 1] - fmin[Ey - 1] * dl + fmin[Hx - 1] * Z0 * dl + Jsmin[1] * dl * Z0 / 2.0 - Msmin[0] * dl / 2.0;
      b_max = b_max * (double)hmax[Ey - 1] - fmax[Ey - 1] * dl - fmax[Hx - 1] * Z0 * dl + Jsmax[1] * dl * Z0 / 2.0 + Msmax[0] * dl / 2.0;
    }
  }
  /* complete the connection */
  *C = a_max - *C;
  *A = a_min - *A;
  *D = b_max - *D;
  *B = b_min - *B;
  if (output_number != 0) {
    *output_face_number += 1;
    double* o1 = &h->face_out[((size_t)(*output_face_number - 1) * 2 + 0) * 6];
    double* o2 = &h->face_out[((size_t)(*output_face_number - 1) * 2 + 1) * 6];
    /* side 1: positive = post-connect, negative = pre-connect;
     * side 2: positive = pre-connect, negative = post-connect. */
    double ap1 = *A, an1 = a_in, bp1 = *B, bn1 = b_in;
    double ap2 = c_in, an2 = *C, bp2 = d_in, bn2 = *D;
    double* o[2] = {o1, o2};
    double ap[2] = {ap1, ap2}, an[2] = {an1, an2}, bp[2] = {bp1, bp2}, bn[2] = {bn1, bn2};
    for (int s = 0; s < 2; s++) {
      double* f = o[s];
      if (axis == 0) { /* connect_normal_to_x.F90:334-355 */
        f[0] = 0.0;
        f[1] = -(ap[s] + an[s]) / dl;
        f[2] = -(bp[s] + bn[s]) / dl;
        f[3] = 0.0;
        f[4] = (bp[s] - bn[s]) / (Z0 * dl);
        f[5] = -(ap[s] - an[s]) / (Z0 * dl);
      } else if (axis == 1) { /* connect_normal_to_y.F90:342-362 */
        f[0] = -(ap[s] + an[s]) / dl;
        f[1] = 0.0;
        f[2] = -(bp[s] + bn[s]) / dl;
        f[3] = -(bp[s] - bn[s]) / (Z0 * dl);
        f[4] = 0.0;
        f[5] = (ap[s] - an[s]) / (Z0 * dl);
      } else { /* connect_normal_to_z.F90:337-361 */
        f[0] = -(ap[s] + an[s]) / dl;
        f[1] = -(bp[s] + bn[s]) / dl;
        f[2] = 0.0;
        f[3] = (bp[s] - bn[s]) / (Z0 * dl);
        f[4] = -(ap[s] - an[s]) / (Z0 * dl);
        f[5] = 0.0;
      }
    }
  }
  return TLM_OK;
}

/* Host-path faces in connect()'s scan order.  mode 0 numbers them (and
 * counts), 1 lists (cx, cy, cz, axis), 2 gathers the four scattered pulses
 * beside each face. */
static void host_face_walk(oracle_handle* h, int mode, int32_t* faces, double* V4) {
  static const int pa[3] = {Vy_xmin, Vx_ymin, Vx_zmin}, pb[3] = {Vz_xmin, Vz_ymin, Vy_zmin};
  static const int qa[3] = {Vy_xmax, Vx_ymax, Vx_zmax}, qb[3] = {Vz_xmax, Vz_ymax, Vy_zmax};
  long long face_number = 0;
  int k = 0;
  for (int cz = 1; cz <= h->nz; cz++)
    for (int cy = 1; cy <= h->ny; cy++)
      for (int cx = 1; cx <= h->nx; cx++)
        for (int axis = 0; axis < 3; axis++) {
          if ((axis == 0 && cx == 1) || (axis == 1 && cy == 1) || (axis == 2 && cz == 1)) continue;
          int code = h->face_code[face_number++];
          if (!code) continue;
          int m = h->f_mat[code - 1];
          int mn = m < 0 ? -m : m;
          if (!mn || !host_smat(h->smat[mn - 1])) continue;
          if (mode == 0) h->hf_of_code[code - 1] = k + 1;
          if (mode == 1 && faces) {
            faces[4 * k] = cx; faces[4 * k + 1] = cy; faces[4 * k + 2] = cz; faces[4 * k + 3] = axis + 1;
          }
          if (mode == 2) {
            const double* v = cellV(h, cx, cy, cz);
            const double* w = cellV(h, cx - (axis == 0), cy - (axis == 1), cz - (axis == 2));
            V4[4 * k] = v[pa[axis]]; V4[4 * k + 1] = w[qa[axis]];
            V4[4 * k + 2] = v[pb[axis]]; V4[4 * k + 3] = w[qb[axis]];
          }
          k++;
        }
  if (mode == 0) h->n_host_faces = k;
}

static int host_face_numbers(oracle_handle* h) {
  if (h->hf_of_code) return TLM_OK;
  if (!h->face_code) return fail(h, TLM_ERR_STATE, "face codes not set");
  for (int i = 0; i < h->n_special_faces; i++) {
    int m = h->f_mat[i] < 0 ? -h->f_mat[i] : h->f_mat[i];
    if (m > h->n_smat) return fail(h, TLM_ERR_ARG, "surface material number out of range");
  }
  h->hf_of_code = (int32_t*)calloc((size_t)(h->n_special_faces ? h->n_special_faces : 1), sizeof(int32_t));
  if (!h->hf_of_code) return fail(h, TLM_ERR_NOMEM, "host face numbers");
  host_face_walk(h, 0, NULL, NULL);
  free(h->hf_tot);
  h->hf_tot = (double*)calloc((size_t)(h->n_host_faces ? h->n_host_faces : 1) * 4, sizeof(double));
  h->hf_ready = 0;
  return h->hf_tot ? TLM_OK : fail(h, TLM_ERR_NOMEM, "host face voltages");
}

int oracle_host_faces(oracle_handle* h, int32_t* n_host_faces, int32_t* faces) {
  if (!h || !n_host_faces) return TLM_ERR_ARG;
  int rc;
  if ((rc = host_face_numbers(h))) return rc;
  *n_host_faces = h->n_host_faces;
  if (faces) host_face_walk(h, 1, faces, NULL);
  return TLM_OK;
}

int oracle_gather_host_faces(oracle_handle* h, double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  int rc;
  if ((rc = host_face_numbers(h))) return rc;
  host_face_walk(h, 2, NULL, V);
  return TLM_OK;
}

int oracle_scatter_host_faces(oracle_handle* h, const double* T) {
  if (!h || !T) return TLM_ERR_ARG;
  int rc;
  if ((rc = host_face_numbers(h))) return rc;
  memcpy(h->hf_tot, T, (size_t)h->n_host_faces * 4 * sizeof(double));
  h->hf_ready = 1;
  return TLM_OK;
}

/* Faces of z-plane cz in connect()'s scan order: the x-face (cx != 1), the
 * y-face (cy != 1) and the z-face (cz != 1) of every cell of the plane. */
static long long faces_before_plane(const oracle_handle* h, int cz) {
  const long long nx = h->nx, ny = h->ny;
  const long long fxy = (nx - 1) * ny + nx * (ny - 1);
  return (long long)(cz - 1) * fxy + (cz > 1 ? (long long)(cz - 2) * nx * ny : 0);
}

/* Scan-order excitation and output face numbers at the start of each
 * z-plane (connect_special_face counts every special face carrying them). */
static int face_prefixes(oracle_handle* h) {
  if (h->pl_fexc) return TLM_OK;
  int nz = h->nz;
  h->pl_fexc = (int32_t*)calloc((size_t)nz + 1, sizeof(int32_t));
  h->pl_fout = (int32_t*)calloc((size_t)nz + 1, sizeof(int32_t));
  if (!h->pl_fexc || !h->pl_fout) return fail(h, TLM_ERR_NOMEM, "plane counters");
  int ne = 0, no = 0;
  for (int cz = 1; cz <= nz; cz++) {
    h->pl_fexc[cz - 1] = ne;
    h->pl_fout[cz - 1] = no;
    for (long long f = faces_before_plane(h, cz); f < faces_before_plane(h, cz + 1); f++) {
      int code = h->face_code[f];
      if (!code) continue;
      ne += h->f_exc[code - 1] != 0;
      no += h->f_out[code - 1] != 0;
    }
  }
  h->pl_fexc[nz] = ne;
  h->pl_fout[nz] = no;
  return TLM_OK;
}

/* connect() over the faces of z-plane cz (connect.F90:180-214): the x- and
 * y-faces inside the plane and the z-faces to plane cz-1.  A face touches
 * only its own two port pairs, so planes can run concurrently. */
static int connect_plane(oracle_handle* h, int cz) {
  int nx = h->nx, ny = h->ny;
  long long face_number = faces_before_plane(h, cz);
  int excitation_face_number = h->pl_fexc[cz - 1], output_face_number = h->pl_fout[cz - 1];
  {
    for (int cy = 1; cy <= ny; cy++)
      for (int cx = 1; cx <= nx; cx++) {
        double* v = cellV(h, cx, cy, cz);
        if (cx != 1) {
          face_number++;
          double* w = cellV(h, cx - 1, cy, cz);
          int code = h->face_code[face_number - 1];
          if (code == 0) { /* connect_normal_to_x.F90:4-10 */
            double Vy = v[Vy_xmin] + w[Vy_xmax];
            v[Vy_xmin] = Vy - v[Vy_xmin];
            w[Vy_xmax] = Vy - w[Vy_xmax];
            double Vz = v[Vz_xmin] + w[Vz_xmax];
            v[Vz_xmin] = Vz - v[Vz_xmin];
            w[Vz_xmax] = Vz - w[Vz_xmax];
          } else {
            int rc = connect_special_face(h, 0, v, w, code, &excitation_face_number, &output_face_number);
            if (rc) return rc;
          }
        }
        if (cy != 1) {
          face_number++;
          double* w = cellV(h, cx, cy - 1, cz);
          int code = h->face_code[face_number - 1];
          if (code == 0) {
            double Vx = v[Vx_ymin] + w[Vx_ymax];
            v[Vx_ymin] = Vx - v[Vx_ymin];
            w[Vx_ymax] = Vx - w[Vx_ymax];
            double Vz = v[Vz_ymin] + w[Vz_ymax];
            v[Vz_ymin] = Vz - v[Vz_ymin];
            w[Vz_ymax] = Vz - w[Vz_ymax];
          } else {
            int rc = connect_special_face(h, 1, v, w, code, &excitation_face_number, &output_face_number);
            if (rc) return rc;
          }
        }
        if (cz != 1) {
          face_number++;
          double* w = cellV(h, cx, cy, cz - 1);
          int code = h->face_code[face_number - 1];
          if (code == 0) {
            double Vx = v[Vx_zmin] + w[Vx_zmax];
            v[Vx_zmin] = Vx - v[Vx_zmin];
            w[Vx_zmax] = Vx - w[Vx_zmax];
            double Vy = v[Vy_zmin] + w[Vy_zmax];
            v[Vy_zmin] = Vy - v[Vy_zmin];
            w[Vy_zmax] = Vy - w[Vy_zmax];
          } else {
            int rc = connect_special_face(h, 2, v, w, code, &excitation_face_number, &output_face_number);
            if (rc) return rc;
          }
        }
      }
  }
  return TLM_OK;
}

int oracle_connect(oracle_handle* h, int32_t timestep) {
  if (!h->face_code) return fail(h, TLM_ERR_STATE, "face codes not set");
  int rc;
  if ((rc = face_prefixes(h))) return rc;
  if ((rc = host_face_numbers(h))) return rc;
  if (h->n_host_faces && !h->hf_ready)
    return fail(h, TLM_ERR_STATE, "host-path faces need tlm_scatter_host_faces between scatter and connect");
  h->hf_ready = 0;
  int err = TLM_OK;
#pragma omp parallel for schedule(static) num_threads(h->threads)
  for (int cz = 1; cz <= h->nz; cz++) {
    int r = connect_plane(h, cz);
    if (r) {
#pragma omp atomic write
      err = r;
    }
  }
  if (err) return err;
  /* face_output_points (output.F90:232-272): this connect's rows, one block
   * per step, kept for reading */
  if (h->n_face_out_rows) {
    size_t w = (size_t)h->n_face_out_rows * 12;
    if (!h->fring) h->fring = (double*)malloc((size_t)h->cap * w * sizeof(double));
    if (h->fring_count == 0) h->fring_first = timestep;
    if (h->fring_count < h->cap) {
      memcpy(h->fring + (size_t)h->fring_count * w, h->face_out, w * sizeof(double));
      h->fring_count++;
    }
    double time = (double)(timestep - 1) * h->g.dt + h->g.dt / 2.0;
    if (h->n_fo_items) accumulate_frequency_outputs(h, 2, h->face_out, timestep, time);
    if (h->n_ff) far_field_currents(h, h->face_out, time);
    if (h->rcs_on) rcs_currents(h, h->face_out, time);
    if (h->n_fd) record_field_outputs(h, 2, h->face_out, timestep);
  }
  return TLM_OK;
}

/* outer_boundary.F90:71-131 (SEQ: rank 0 owns both z faces). */
int oracle_outer_boundary(oracle_handle* h) {
  int nx = h->nx, ny = h->ny, nz = h->nz;
  const tlm_grid* g = &h->g;
  for (int cz = 1; cz <= nz; cz++)
    for (int cy = 1; cy <= ny; cy++) {
      double* v = cellV(h, 1, cy, cz);
      v[Vy_xmin] = v[Vy_xmin] * g->r_xmin;
      v[Vz_xmin] = v[Vz_xmin] * g->r_xmin;
      v = cellV(h, nx, cy, cz);
      v[Vy_xmax] = v[Vy_xmax] * g->r_xmax;
      v[Vz_xmax] = v[Vz_xmax] * g->r_xmax;
    }
  for (int cz = 1; cz <= nz; cz++)
    for (int cx = 1; cx <= nx; cx++) {
      double* v = cellV(h, cx, 1, cz);
      v[Vx_ymin] = v[Vx_ymin] * g->r_ymin;
      v[Vz_ymin] = v[Vz_ymin] * g->r_ymin;
      v = cellV(h, cx, ny, cz);
      v[Vx_ymax] = v[Vx_ymax] * g->r_ymax;
      v[Vz_ymax] = v[Vz_ymax] * g->r_ymax;
    }
  for (int cy = 1; cy <= ny; cy++)
    for (int cx = 1; cx <= nx; cx++) {
      double* v = cellV(h, cx, cy, 1);
      v[Vx_zmin] = v[Vx_zmin] * g->r_zmin;
      v[Vy_zmin] = v[Vy_zmin] * g->r_zmin;
    }
  for (int cy = 1; cy <= ny; cy++)
    for (int cx = 1; cx <= nx; cx++) {
      double* v = cellV(h, cx, cy, nz);
      v[Vx_zmax] = v[Vx_zmax] * g->r_zmax;
      v[Vy_zmax] = v[Vy_zmax] * g->r_zmax;
    }
  return TLM_OK;
}

int oracle_step(oracle_handle* h, int32_t first_timestep, int32_t n_steps) {
  if (h->face_code && h->smat) {
    int rc = host_face_numbers(h);
    if (rc) return rc;
    if (h->n_host_faces) return fail(h, TLM_ERR_STATE, "host-path faces need the split calls (tlm_scatter, tlm_connect)");
  }
  if (h->c_cable)
    for (int32_t i = 0; i < h->n_special; i++)
      if (h->c_cable[i]) return fail(h, TLM_ERR_STATE, "host-path cells need the split calls (tlm_scatter, tlm_connect)");
  for (int32_t t = first_timestep; t < first_timestep + n_steps; t++) {
    int rc;
    if ((rc = oracle_excitation(h, t))) return rc;
    if ((rc = oracle_scatter(h, t))) return rc;
    if ((rc = oracle_connect(h, t))) return rc;
    if ((rc = oracle_outer_boundary(h))) return rc;
  }
  return TLM_OK;
}

int oracle_synchronize(oracle_handle* h) { (void)h; return TLM_OK; }

/* Threads over z-slabs for scatter and connect (1: the sequential build).
 * Every cell and face is computed by the same operations whatever the
 * count, so the results are identical. */
int oracle_set_threads(oracle_handle* h, int32_t threads) {
  if (!h || threads < 1) return TLM_ERR_ARG;
  h->threads = threads;
  return TLM_OK;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int oracle_set_output_capacity(oracle_handle* h, int32_t steps) {
  if (steps < 1) return fail(h, TLM_ERR_ARG, "capacity must be positive");
  free(h->ring);
  h->ring = NULL;
  h->cap = steps;
  h->ring_count = 0;
  return TLM_OK;
}

int oracle_read_cell_outputs(oracle_handle* h, double* out, int32_t max_steps,
                             int32_t* n_read, int32_t* first_step) {
  int rows = h->n_out_rows;
  int n = h->ring_count < max_steps ? h->ring_count : max_steps;
  if (n > 0) memcpy(out, h->ring, (size_t)n * rows * 6 * sizeof(double));
  if (n_read) *n_read = n;
  if (first_step) *first_step = h->ring_first;
  if (n < h->ring_count) {
    memmove(h->ring, h->ring + (size_t)n * rows * 6, (size_t)(h->ring_count - n) * rows * 6 * sizeof(double));
    h->ring_first += n;
  }
  h->ring_count -= n;
  return TLM_OK;
}

int oracle_set_frequency_outputs(oracle_handle* h, int32_t n, const tlm_frequency_output* outs, const int32_t* rows,
                                 const int32_t* sides, const int32_t* normals) {
  if (n < 0 || (n && !outs)) return TLM_ERR_ARG;
  int items = 0;
  for (int i = 0; i < n; i++) {
    const tlm_frequency_output* f = &outs[i];
    if ((f->kind != 1 && f->kind != 2) || f->n_items < 0 || f->timestep_interval < 1 || f->field_component < 1 ||
        f->field_component > (f->kind == 1 ? 6 : 9))
      return fail(h, TLM_ERR_ARG, "frequency output: bad kind, component or window");
    items += f->n_items;
  }
  if (items && (!rows || !sides || !normals)) return TLM_ERR_ARG;
  free(h->fo); free(h->fo_out); free(h->fo_row); free(h->fo_side); free(h->fo_norm); free(h->fo_acc);
  h->n_fo = n;
  h->n_fo_items = items;
  h->fo = (tlm_frequency_output*)malloc((size_t)(n ? n : 1) * sizeof(tlm_frequency_output));
  if (n) memcpy(h->fo, outs, (size_t)n * sizeof(tlm_frequency_output));
  h->fo_out = (int32_t*)malloc((size_t)(items ? items : 1) * sizeof(int32_t));
  h->fo_row = dup_i32(rows, (size_t)items);
  h->fo_side = dup_i32(sides, (size_t)items);
  h->fo_norm = dup_i32(normals, (size_t)items);
  h->fo_acc = (double*)calloc((size_t)(items ? items : 1) * 2, sizeof(double));
  int k = 0;
  for (int o = 0; o < n; o++) {
    int nrows = outs[o].kind == 1 ? h->n_out_rows : h->n_face_out_rows;
    for (int i = 0; i < outs[o].n_items; i++, k++) {
      h->fo_out[k] = o;
      if (h->fo_row[k] < 1 || h->fo_row[k] > nrows) return fail(h, TLM_ERR_ARG, "frequency output: row out of range");
      if (outs[o].kind == 2 && (h->fo_side[k] < 1 || h->fo_side[k] > 2 || h->fo_norm[k] == 0 ||
                                h->fo_norm[k] < -3 || h->fo_norm[k] > 3))
        return fail(h, TLM_ERR_ARG, "frequency output: bad face side or normal");
    }
  }
  return TLM_OK;
}

int oracle_read_frequency_output(oracle_handle* h, int32_t k, double* re_im) {
  if (k < 0 || k >= h->n_fo || !re_im) return TLM_ERR_ARG;
  size_t off = 0;
  for (int i = 0; i < k; i++) off += (size_t)h->fo[i].n_items;
  memcpy(re_im, h->fo_acc + off * 2, (size_t)h->fo[k].n_items * 2 * sizeof(double));
  return TLM_OK;
}

int oracle_set_field_outputs(oracle_handle* h, int32_t n, const tlm_field_output* outs, const int32_t* rows,
                             const int32_t* sides, const int32_t* normals) {
  if (n < 0 || (n && !outs)) return TLM_ERR_ARG;
  size_t items = 0;
  for (int i = 0; i < n; i++) {
    const tlm_field_output* f = &outs[i];
    int c = f->field_component;
    int surf_c = (c >= 1 && c <= 9) || (c >= 13 && c <= 15);
    int ok = (f->kind == 1 && c >= 1 && c <= 6 && !f->both_sides) ||
             (f->kind == 2 && surf_c && (!f->both_sides || (c >= 7 && c <= 9) || c == 15));
    if (!ok || f->n_items < 0 || f->timestep_interval < 1 || (f->both_sides != 0 && f->both_sides != 1))
      return fail(h, TLM_ERR_ARG, "field output: bad kind, component or window");
    items += (size_t)f->n_items;
  }
  if (items && (!rows || !sides || !normals)) return TLM_ERR_ARG;
  size_t k = 0;
  for (int i = 0; i < n; i++) {
    int nrows = outs[i].kind == 1 ? h->n_out_rows : h->n_face_out_rows;
    for (int j = 0; j < outs[i].n_items; j++, k++) {
      if (rows[k] < 1 || rows[k] > nrows) return fail(h, TLM_ERR_ARG, "field output: row out of range");
      if (outs[i].kind == 2 && (sides[k] < 1 || sides[k] > 2 || normals[k] == 0 || normals[k] < -3 || normals[k] > 3))
        return fail(h, TLM_ERR_ARG, "field output: bad face side or normal");
    }
  }
  for (int i = 0; i < h->n_fd; i++) free(h->fd_frames[i]);
  free(h->fd); free(h->fd_row); free(h->fd_side); free(h->fd_norm); free(h->fd_frames); free(h->fd_nframes);
  free(h->fd_capframes); free(h->fd_first);
  h->n_fd = n;
  h->fd = (tlm_field_output*)malloc((size_t)(n ? n : 1) * sizeof(tlm_field_output));
  if (n) memcpy(h->fd, outs, (size_t)n * sizeof(tlm_field_output));
  h->fd_row = dup_i32(rows, items);
  h->fd_side = dup_i32(sides, items);
  h->fd_norm = dup_i32(normals, items);
  h->fd_frames = (double**)calloc((size_t)(n ? n : 1), sizeof(double*));
  h->fd_nframes = (int*)calloc((size_t)(n ? n : 1), sizeof(int));
  h->fd_capframes = (int*)calloc((size_t)(n ? n : 1), sizeof(int));
  h->fd_first = (int*)calloc((size_t)(n ? n : 1), sizeof(int));
  return TLM_OK;
}

int oracle_read_field_output(oracle_handle* h, int32_t k, double* values, int32_t max_frames, int32_t* n_frames,
                             int32_t* first_timestep) {
  if (k < 0 || k >= h->n_fd || max_frames < 0 || !n_frames || !first_timestep || (max_frames && !values))
    return TLM_ERR_ARG;
  const tlm_field_output* f = &h->fd[k];
  int n = max_frames < h->fd_nframes[k] ? max_frames : h->fd_nframes[k];
  size_t items = (size_t)f->n_items;
  if (n && items) memcpy(values, h->fd_frames[k], (size_t)n * items * sizeof(double));
  *n_frames = n;
  *first_timestep = h->fd_first[k];
  if (n && items)
    memmove(h->fd_frames[k], h->fd_frames[k] + (size_t)n * items,
            (size_t)(h->fd_nframes[k] - n) * items * sizeof(double));
  h->fd_nframes[k] -= n;
  h->fd_first[k] += n * f->timestep_interval;
  return TLM_OK;
}

int oracle_set_far_field_surfaces(oracle_handle* h, int32_t n, const tlm_far_field_surface* surfaces,
                                  const int32_t* rows, const int32_t* sides, const int32_t* normals,
                                  const double* centres) {
  if (n < 0 || (n && !surfaces)) return TLM_ERR_ARG;
  size_t faces = 0;
  for (int i = 0; i < n; i++) {
    const tlm_far_field_surface* f = &surfaces[i];
    if (f->n_faces < 0 || f->n_theta < 1 || f->n_phi < 1 || (f->dim != 2 && f->dim != 3) ||
        (f->dim == 2 && f->n_theta != 1))
      return fail(h, TLM_ERR_ARG, "far-field surface: bad dimension or angles");
    faces += (size_t)f->n_faces;
  }
  if (faces && (!rows || !sides || !normals || !centres)) return TLM_ERR_ARG;
  for (size_t i = 0; i < faces; i++) {
    if (rows[i] < 1 || rows[i] > h->n_face_out_rows) return fail(h, TLM_ERR_ARG, "far-field surface: row out of range");
    if (sides[i] < 1 || sides[i] > 2 || normals[i] == 0 || normals[i] < -3 || normals[i] > 3)
      return fail(h, TLM_ERR_ARG, "far-field surface: bad face side or normal");
  }
  free(h->ff); free(h->ff_row); free(h->ff_side); free(h->ff_norm); free(h->ff_centre); free(h->ff_J); free(h->ff_M);
  h->n_ff = n;
  h->ff = (tlm_far_field_surface*)malloc((size_t)(n ? n : 1) * sizeof(tlm_far_field_surface));
  if (n) memcpy(h->ff, surfaces, (size_t)n * sizeof(tlm_far_field_surface));
  h->ff_row = dup_i32(rows, faces);
  h->ff_side = dup_i32(sides, faces);
  h->ff_norm = dup_i32(normals, faces);
  h->ff_centre = (double*)malloc((faces ? faces : 1) * 3 * sizeof(double));
  if (faces) memcpy(h->ff_centre, centres, faces * 3 * sizeof(double));
  h->ff_J = (double*)calloc((faces ? faces : 1) * 6, sizeof(double));
  h->ff_M = (double*)calloc((faces ? faces : 1) * 6, sizeof(double));
  return TLM_OK;
}

/* write_far_field_surfaces (far_field_surfaces.F90:370-712): the N and L
   integrals over the faces for each angle, summed in face order, and the
   far field at r0 = 1 m; complex arithmetic written out as the Fortran
   complex*16 operations it restates. */
int oracle_read_far_field(oracle_handle* h, int32_t k, double* e) {
  if (k < 0 || k >= h->n_ff || !e) return TLM_ERR_ARG;
  const tlm_far_field_surface* f = &h->ff[k];
  size_t f0 = 0;
  for (int i = 0; i < k; i++) f0 += (size_t)h->ff[i].n_faces;
  double dA = f->dim == 2 ? h->g.dl : h->g.dl * h->g.dl;
  /* jbeta = j*cmplx(beta) (:431, :561): CMPLX without a kind is a
     single-precision complex, so beta enters rounded to a float */
  double beta = (double)(float)(2.0 * PI * f->frequency / C0);
  double r0 = 1.0;
  double cb = cos(beta * r0), sb = sin(beta * r0); /* exp(j beta r0) */
  double p2r, p2i;
  if (f->dim == 2) {
    /* sqrt(j beta / (8 pi r0)): the principal root of j y, y > 0, is
       sqrt(y/2) (1 + j) */
    double rt = sqrt(0.5 * (beta / (8.0 * PI * r0)));
    p2r = rt * cb - rt * sb;
    p2i = rt * sb + rt * cb;
  } else {
    p2r = (0.0 * cb - beta * sb) / (4.0 * PI * r0);
    p2i = (0.0 * sb + beta * cb) / (4.0 * PI * r0);
  }
  int count = 0;
  for (int it = 0; it < f->n_theta; it++) {
    double theta = f->theta_min + (double)it * f->theta_step;
    for (int ip = 0; ip < f->n_phi; ip++, count++) {
      double phi = f->phi_min + (double)ip * f->phi_step;
      double r[3] = {r0 * sin(theta) * cos(phi), r0 * sin(theta) * sin(phi), r0 * cos(theta)};
      double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
      double a[8] = {0, 0, 0, 0, 0, 0, 0, 0}; /* N_theta|N_z, N_phi, L_theta|L_z, L_phi */
      for (int i = 0; i < f->n_faces; i++) {
        size_t q = f0 + (size_t)i;
        const double* rp = &h->ff_centre[3 * q];
        double x = f->dim == 2 ? r[0] * rp[0] + r[1] * rp[1] : r[0] * rp[0] + r[1] * rp[1] + r[2] * rp[2];
        double pr = cos(beta * x), pi = sin(beta * x); /* exp(j beta x) */
        const double* J = &h->ff_J[q * 6];
        const double* M = &h->ff_M[q * 6];
        double t[8];
        if (f->dim == 2) {
          t[0] = J[4]; t[1] = J[5];
          t[2] = -J[0] * sp + J[2] * cp; t[3] = -J[1] * sp + J[3] * cp;
          t[4] = M[4]; t[5] = M[5];
          t[6] = -M[0] * sp + M[2] * cp; t[7] = -M[1] * sp + M[3] * cp;
        } else {
          t[0] = J[0] * ct * cp + J[2] * ct * sp - J[4] * st; t[1] = J[1] * ct * cp + J[3] * ct * sp - J[5] * st;
          t[2] = -J[0] * sp + J[2] * cp; t[3] = -J[1] * sp + J[3] * cp;
          t[4] = M[0] * ct * cp + M[2] * ct * sp - M[4] * st; t[5] = M[1] * ct * cp + M[3] * ct * sp - M[5] * st;
          t[6] = -M[0] * sp + M[2] * cp; t[7] = -M[1] * sp + M[3] * cp;
        }
        for (int c = 0; c < 4; c++) {
          double zr = t[2 * c] * pr - t[2 * c + 1] * pi, zi = t[2 * c] * pi + t[2 * c + 1] * pr;
          a[2 * c] = a[2 * c] + zr * dA;
          a[2 * c + 1] = a[2 * c + 1] + zi * dA;
        }
      }
      double qr, qi, ur, ui;
      if (f->dim == 2) {
        qr = a[6] - Z0 * a[0]; qi = a[7] - Z0 * a[1]; /* L_phi - Z0 N_z */
        ur = a[4] + Z0 * a[2]; ui = a[5] + Z0 * a[3]; /* L_z + Z0 N_phi */
      } else {
        qr = a[6] + Z0 * a[0]; qi = a[7] + Z0 * a[1]; /* L_phi + Z0 N_theta */
        ur = a[4] - Z0 * a[2]; ui = a[5] - Z0 * a[3]; /* L_theta - Z0 N_phi */
      }
      e[4 * count] = -(p2r * qr - p2i * qi);
      e[4 * count + 1] = -(p2r * qi + p2i * qr);
      e[4 * count + 2] = p2r * ur - p2i * ui;
      e[4 * count + 3] = p2r * ui + p2i * ur;
    }
  }
  return TLM_OK;
}

int oracle_set_rcs_surface(oracle_handle* h, const tlm_rcs_surface* s, const int32_t* rows, const int32_t* sides,
                           const int32_t* normals, const double* centres) {
  free(h->rcs_row); free(h->rcs_side); free(h->rcs_norm); free(h->rcs_centre); free(h->rcs_E);
  h->rcs_row = h->rcs_side = h->rcs_norm = NULL;
  h->rcs_centre = h->rcs_E = NULL;
  h->rcs_on = 0;
  if (!s) return TLM_OK;
  if (s->n_faces < 0 || s->n_points < 1 || s->points_per_cell < 1 ||
      (s->n_faces && (!rows || !sides || !normals || !centres)))
    return fail(h, TLM_ERR_ARG, "RCS surface: bad faces or far-field points");
  for (int i = 0; i < s->n_faces; i++) {
    if (rows[i] < 1 || rows[i] > h->n_face_out_rows) return fail(h, TLM_ERR_ARG, "RCS surface: row out of range");
    if (sides[i] < 1 || sides[i] > 2 || normals[i] == 0 || normals[i] < -3 || normals[i] > 3)
      return fail(h, TLM_ERR_ARG, "RCS surface: bad face side or normal");
  }
  h->rcs_on = 1;
  h->rcs = *s;
  h->rcs_row = dup_i32(rows, (size_t)s->n_faces);
  h->rcs_side = dup_i32(sides, (size_t)s->n_faces);
  h->rcs_norm = dup_i32(normals, (size_t)s->n_faces);
  h->rcs_centre = (double*)malloc((size_t)(s->n_faces ? s->n_faces : 1) * 3 * sizeof(double));
  if (s->n_faces) memcpy(h->rcs_centre, centres, (size_t)s->n_faces * 3 * sizeof(double));
  h->rcs_E = (double*)calloc((size_t)s->n_points * 4, sizeof(double));
  return TLM_OK;
}

int oracle_read_rcs(oracle_handle* h, double* out) {
  if (!h->rcs_on || !out) return TLM_ERR_ARG;
  memcpy(out, h->rcs_E, (size_t)h->rcs.n_points * 4 * sizeof(double));
  return TLM_OK;
}

int oracle_read_dft(oracle_handle* h, double* re_im) {
  size_t n = (size_t)h->n_points * h->n_freq * 2;
  if (n) memcpy(re_im, h->dft, n * sizeof(double));
  return TLM_OK;
}

int oracle_read_face_outputs(oracle_handle* h, double* out, int32_t max_steps, int32_t* n_read,
                             int32_t* first_step) {
  size_t w = (size_t)h->n_face_out_rows * 12;
  int n = h->fring_count < max_steps ? h->fring_count : max_steps;
  if (n > 0 && w) memcpy(out, h->fring, (size_t)n * w * sizeof(double));
  if (n_read) *n_read = n;
  if (first_step) *first_step = h->fring_first;
  if (n < h->fring_count) {
    memmove(h->fring, h->fring + (size_t)n * w, (size_t)(h->fring_count - n) * w * sizeof(double));
    h->fring_first += n;
  }
  h->fring_count -= n;
  return TLM_OK;
}

/* Schedule and measurement hooks of the engine ABI; the oracle has no
 * kernels to schedule or time. */
int oracle_set_schedule(oracle_handle* h, int32_t variant, int32_t rows, int32_t zc, int32_t xcd_remap,
                        int32_t mixed) {
  (void)h; (void)variant; (void)rows; (void)zc; (void)xcd_remap; (void)mixed;
  return TLM_OK;
}
int oracle_set_profiling(oracle_handle* h, int32_t enable) { (void)h; (void)enable; return TLM_OK; }

/* The oracle stores V in the reference's scan order whatever the layout:
 * one segment. */
int oracle_set_layout(oracle_handle* h, int32_t layout) {
  if (layout != TLM_LAYOUT_X_SEGMENTS && layout != TLM_LAYOUT_ROWS) return fail(h, TLM_ERR_ARG, "bad layout");
  h->layout = layout;
  return TLM_OK;
}
/* Host-path (cable) cells in scan order (x fastest), 1-based coordinates. */
static int host_cell_scan(oracle_handle* h, int32_t* n, int32_t* cells, double* out, const double* in) {
  if (!h->cell_code) return fail(h, TLM_ERR_STATE, "cell codes not set");
  int32_t k = 0, cell_number = 0;
  for (int cz = 1; cz <= h->nz; cz++)
    for (int cy = 1; cy <= h->ny; cy++)
      for (int cx = 1; cx <= h->nx; cx++) {
        int code = h->cell_code[cell_number++];
        if (code == 0 || !h->c_cable[code - 1]) continue;
        if (cells) { cells[3 * k] = cx; cells[3 * k + 1] = cy; cells[3 * k + 2] = cz; }
        double* v = cellV(h, cx, cy, cz);
        for (int p = 0; p < 12; p++) {
          if (out) out[12 * k + p] = v[p + 1];
          if (in) v[p + 1] = in[12 * k + p];
        }
        k++;
      }
  if (n) *n = k;
  return TLM_OK;
}

int oracle_host_cells(oracle_handle* h, int32_t* n_host, int32_t* cells) {
  if (!h || !n_host) return TLM_ERR_ARG;
  return host_cell_scan(h, n_host, cells, NULL, NULL);
}

int oracle_gather_host_cells(oracle_handle* h, double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  return host_cell_scan(h, NULL, NULL, V, NULL);
}

int oracle_scatter_host_cells(oracle_handle* h, const double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  return host_cell_scan(h, NULL, NULL, NULL, V);
}

int oracle_layout_segments(oracle_handle* h, int32_t* n_segments, int32_t* x0) {
  if (!n_segments) return TLM_ERR_ARG;
  *n_segments = 1;
  if (x0) { x0[0] = 0; x0[1] = h->nx; }
  return TLM_OK;
}

int oracle_get_schedule(const oracle_handle* h, int32_t* variant, int32_t* rows, int32_t* zc, int32_t* xcd_remap,
                        int32_t* mixed, int32_t* streams) {
  (void)h;
  if (variant) *variant = 3;
  if (rows) *rows = 1;
  if (zc) *zc = 1;
  if (xcd_remap) *xcd_remap = 0;
  if (mixed) *mixed = 1;
  if (streams) *streams = 1;
  return TLM_OK;
}

int oracle_set_streams(oracle_handle* h, int32_t streams) {
  if (streams != 1 && streams != 2) return fail(h, TLM_ERR_ARG, "streams must be 1 or 2");
  return TLM_OK;
}

/* The oracle has one schedule: the autotuner just steps, as many steps as
 * the engine's candidates take (8, or 14 with TLM_LAYOUT_ROWS). */
int oracle_autotune(oracle_handle* h, int32_t first_timestep, int32_t steps_per_candidate, int32_t* steps_done) {
  int n = (h->layout == TLM_LAYOUT_ROWS ? 14 : 8) * (steps_per_candidate + 1);
  int rc = oracle_step(h, first_timestep, n);
  if (!rc && steps_done) *steps_done = n;
  return rc;
}

int oracle_set_small_to_zero(oracle_handle* h, int32_t n_small_to_zero, double small_to_zero_value) {
  if (n_small_to_zero < 0 || !(small_to_zero_value >= 0.0)) return fail(h, TLM_ERR_ARG, "bad small_to_zero data");
  h->stz_n = n_small_to_zero;
  h->stz_value = small_to_zero_value;
  return TLM_OK;
}
int oracle_read_kernel_times(oracle_handle* h, int64_t* launches, double* ms) {
  (void)h;
  for (int k = 0; k < TLM_N_KCLASS; k++) { launches[k] = 0; ms[k] = 0.0; }
  return TLM_OK;
}
int oracle_kernel_bytes(oracle_handle* h, int64_t* cells, int64_t* bytes) {
  (void)h;
  for (int k = 0; k < TLM_N_KCLASS; k++) { cells[k] = 0; bytes[k] = 0; }
  return TLM_OK;
}

/* Test hook: the per-step filter sequence of a dispersive cell component
 * (timeshift, history-only evaluate, evaluate with the input) applied to
 * G[0..n-1]; out[i] is the filter output after step i. */
int oracle_zfilter_sequence(const double* a, const double* b, int32_t order, const double* G,
                            int32_t n, double* out) {
  double w[NW] = {0};
  if (order < 0 || order > TLM_MAX_FILTER_ORDER) return TLM_ERR_ARG;
  for (int i = 0; i < n; i++) {
    timeshift_Zfilter(w, order);
    (void)evaluate_Zfilter(a, b, w, order, 0.0);
    out[i] = evaluate_Zfilter(a, b, w, order, G[i]);
  }
  return TLM_OK;
}
