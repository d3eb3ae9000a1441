/* tlmbatch.h -- C ABI of the batch engine in libtlmscn.so: many resident
 * solves stepped together (continuous batching), for serving loops.
 *
 * No counterpart in the reference, which runs one solve per process.  A
 * serving loop keeps many configured handles (include/tlmscn.h) resident
 * in HBM and steps them together: one launch per kernel class per step over
 * all of them, replayed from a HIP graph of K steps.  A solve gives the same
 * results alone (tlm_step) and in any batch.
 */
#ifndef TLMBATCH_H
#define TLMBATCH_H

#include "tlmscn.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tlm_batch tlm_batch;

/* A batch of solves of one numerical form, admitting solves while their
 * device bytes fit hbm_budget_bytes. */
int tlm_batch_create(int32_t form, int64_t hbm_budget_bytes, tlm_batch** out);
/* Frees the batch's own memory; handles still inside stay valid (destroy
 * them separately). */
void tlm_batch_destroy(tlm_batch* b);
const char* tlm_batch_last_error(const tlm_batch* b);
/* Admit a configured, not yet stepped solve for steps first_timestep ..
 * first_timestep+n_steps-1.  TLM_ERR_BUDGET: it does not fit now (keep it
 * queued and retry after tlm_batch_collect).  The handle must not be used
 * until collected.  Host-path cells and faces are refused
 * (TLM_ERR_UNSUPPORTED): the host exchanges them between the split calls
 * of the single-solve path. */
int tlm_batch_add(tlm_batch* b, tlm_handle* h, int32_t first_timestep, int32_t n_steps);
/* Enqueue n_steps steps of every unfinished solve (solves that reach their
 * last step earlier stop there) and return at once: the host can build the
 * next arrivals meanwhile.  A change of the set of solves (add, collect)
 * takes effect at the next run. */
int tlm_batch_run(tlm_batch* b, int32_t n_steps);
/* Wait until the enqueued steps are done. */
int tlm_batch_synchronize(tlm_batch* b);
/* Wait for the enqueued steps, then hand back up to max_done finished
 * solves.  Each is then as after
 * tlm_step over its range: tlm_read_cell_outputs, tlm_read_face_outputs
 * (the batch recorded the face outputs of every connect it pulled; the
 * last connect's follow when the handle applies it), tlm_read_dft and
 * tlm_download_links work on it. */
int tlm_batch_collect(tlm_batch* b, tlm_handle** done, int32_t max_done, int32_t* n_done);
int tlm_batch_info(const tlm_batch* b, int32_t* n_solves, int32_t* n_unfinished, int64_t* resident_bytes);
/* Steps per graph replay (default 32). */
int tlm_batch_set_graph_steps(tlm_batch* b, int32_t k);
/* Kernels of a batch step: mixed = 1 LDS-staged mixed-line kernel, 0 lane
 * slots, -1 the default (LDS-staged); streams = 2 (default) runs the PML,
 * material, generic and PML-material kernels on a second stream
 * concurrently with the bulk-line and mixed-line kernels, 1 runs every
 * kernel in turn.  Results do not depend on either. */
int tlm_batch_set_kernels(tlm_batch* b, int32_t mixed, int32_t streams);
/* Time candidate graph lengths (4, 16, 64 steps) with each mixed-line
 * kernel of the form (the extended form has the LDS-staged one only) on
 * the current solves (these are real steps: *steps_done of them
 * were run) and keep the fastest. */
int tlm_batch_autotune(tlm_batch* b, int32_t steps_per_candidate, int32_t* best_k, int32_t* steps_done);

#ifdef __cplusplus
}
#endif

#endif /* TLMBATCH_H */
