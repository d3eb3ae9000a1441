/*
 * scn_oracle.h -- CPU oracle for the SCN scatter/connect hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of
 * GGI_TLM v2.1.7's per-timestep loop body, following the Fortran operation
 * order line by line so that its results can be compared bit for bit with
 * the device path (z-slab threads give the same bits as one thread).  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it; the product library never does.
 *
 * It exports the same entry points as include/tlmscn.h with the prefix
 * oracle_ instead of tlm_, taking the same table structs, so a test drives
 * both through one code path.
 */
#ifndef SCN_ORACLE_H
#define SCN_ORACLE_H

#include "../include/tlmscn.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_handle oracle_handle;

int oracle_create(const tlm_grid* grid, int32_t form, oracle_handle** out);
void oracle_destroy(oracle_handle* h);
const char* oracle_last_error(const oracle_handle* h);
int64_t oracle_device_bytes(const oracle_handle* h);
int oracle_set_cell_codes(oracle_handle* h, const int32_t* code, int32_t n_special,
                          const int32_t* to_material, const int32_t* to_pml,
                          const int32_t* to_excitation, const int32_t* to_output,
                          const int32_t* to_cable);
int oracle_set_face_codes(oracle_handle* h, const int32_t* code, int32_t n_special,
                          const int32_t* to_material, const int32_t* to_excitation,
                          const int32_t* to_output, const int32_t* to_cable);
int oracle_set_volume_materials(oracle_handle* h, int32_t n, const tlm_volume_material* m);
int oracle_set_surface_materials(oracle_handle* h, int32_t n, const int32_t* type);
int oracle_set_pml(oracle_handle* h, int32_t n_params, const tlm_pml_parameter* p,
                   int32_t n_cells, const int32_t* cell_param);
int oracle_set_excitation_functions(oracle_handle* h, int32_t n_functions,
                                    const double* value, const double* value_face);
int oracle_set_point_sources(oracle_handle* h, int32_t n, const tlm_point_source* s,
                             int32_t n_cell_rows, int32_t n_face_rows);
int oracle_set_huygens_surface(oracle_handle* h, int32_t function, double offset_min,
                               const double* Ei, const double* Hi, int32_t n_faces,
                               const tlm_huygens_face* faces);
int oracle_set_output_points(oracle_handle* h, int32_t n, const int32_t* row,
                             const int32_t* component);
int oracle_set_dft_frequencies(oracle_handle* h, int32_t n, const double* freq);
int oracle_upload_links(oracle_handle* h, const double* V);
int oracle_download_links(oracle_handle* h, double* V);
int oracle_download_links_noalloc(oracle_handle* h, double* V);
int oracle_excitation(oracle_handle* h, int32_t timestep);
int oracle_scatter(oracle_handle* h, int32_t timestep);
int oracle_connect(oracle_handle* h, int32_t timestep);
int oracle_outer_boundary(oracle_handle* h);
int oracle_step(oracle_handle* h, int32_t first_timestep, int32_t n_steps);
int oracle_synchronize(oracle_handle* h);
int oracle_set_output_capacity(oracle_handle* h, int32_t steps);
int oracle_read_cell_outputs(oracle_handle* h, double* out, int32_t max_steps,
                             int32_t* n_read, int32_t* first_step);
int oracle_read_dft(oracle_handle* h, double* re_im);
int oracle_set_frequency_outputs(oracle_handle* h, int32_t n, const tlm_frequency_output* outputs,
                                 const int32_t* rows, const int32_t* sides, const int32_t* normals);
int oracle_read_frequency_output(oracle_handle* h, int32_t k, double* re_im);
int oracle_set_field_outputs(oracle_handle* h, int32_t n, const tlm_field_output* outputs, const int32_t* rows,
                             const int32_t* sides, const int32_t* normals);
int oracle_read_field_output(oracle_handle* h, int32_t k, double* values, int32_t max_frames, int32_t* n_frames,
                             int32_t* first_timestep);
int oracle_set_far_field_surfaces(oracle_handle* h, int32_t n, const tlm_far_field_surface* surfaces,
                                  const int32_t* rows, const int32_t* sides, const int32_t* normals,
                                  const double* centres);
int oracle_read_far_field(oracle_handle* h, int32_t k, double* e);
int oracle_set_rcs_surface(oracle_handle* h, const tlm_rcs_surface* surface, const int32_t* rows,
                           const int32_t* sides, const int32_t* normals, const double* centres);
int oracle_read_rcs(oracle_handle* h, double* out);
int oracle_set_schedule(oracle_handle* h, int32_t variant, int32_t rows, int32_t zc, int32_t xcd_remap,
                        int32_t mixed);
int oracle_set_profiling(oracle_handle* h, int32_t enable);
int oracle_set_layout(oracle_handle* h, int32_t layout);
int oracle_set_streams(oracle_handle* h, int32_t streams);
int oracle_get_schedule(const oracle_handle* h, int32_t* variant, int32_t* rows, int32_t* zc, int32_t* xcd_remap,
                        int32_t* mixed, int32_t* streams);
int oracle_layout_segments(oracle_handle* h, int32_t* n_segments, int32_t* x0);
/* host-path (cable) cells: the scatter leaves them incident */
int oracle_host_cells(oracle_handle* h, int32_t* n_host, int32_t* cells);
int oracle_gather_host_cells(oracle_handle* h, double* V);
int oracle_scatter_host_cells(oracle_handle* h, const double* V);
int oracle_host_faces(oracle_handle* h, int32_t* n_host_faces, int32_t* faces);
int oracle_gather_host_faces(oracle_handle* h, double* V);
int oracle_scatter_host_faces(oracle_handle* h, const double* T);
int oracle_autotune(oracle_handle* h, int32_t first_timestep, int32_t steps_per_candidate, int32_t* steps_done);
int oracle_set_small_to_zero(oracle_handle* h, int32_t n_small_to_zero, double small_to_zero_value);
int oracle_read_kernel_times(oracle_handle* h, int64_t* launches, double* ms);
int oracle_kernel_bytes(oracle_handle* h, int64_t* cells, int64_t* bytes);
/* face_output_field(row, side, 1:6) of the last connect, row-major
 * (row, side, component); n_rows = total_number_output_faces. */
int oracle_read_face_outputs(oracle_handle* h, double* out, int32_t max_steps, int32_t* n_read,
                             int32_t* first_step);
/* z-slab threads of scatter and connect; results do not depend on them */
int oracle_set_threads(oracle_handle* h, int32_t threads);
int oracle_max_threads(void);
int oracle_zfilter_sequence(const double* a, const double* b, int32_t order, const double* G,
                            int32_t n, double* out);

#ifdef __cplusplus
}
#endif

#endif
