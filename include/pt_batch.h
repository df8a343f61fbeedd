/* pt_batch.h — many render jobs per submission.
 *
 * The reference renders one view per Renderer and waits on the host after
 * every tile (Renderer.render / renderSection, tracer.js:2242-2309).  A batch
 * takes several contexts (each with its own scene, frame and options, set up
 * through pt_abi.h) and renders one progressive frame of all of them as one
 * graph launch: every job's kernel chain (the same kernels pt_trace issues)
 * is a branch of the graph, so small jobs run side by side on the GPU and the
 * host waits once per frame.  The graph is recorded on the first frame and
 * updated in place on the next ones.
 *
 * No reference counterpart beyond the per-view loop it replaces; results per
 * job are bit-identical to pt_trace on that job's context with the same
 * arguments.  Every call returns PT_OK or a negative PT_ERR_*; messages via
 * pt_batch_last_error. */
#ifndef PT_BATCH_H
#define PT_BATCH_H

#include <stdint.h>

#include "pt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pt_batch pt_batch;

int pt_batch_create(pt_batch** out, int device);
void pt_batch_destroy(pt_batch* b);
const char* pt_batch_last_error(const pt_batch* b);

/* Adds a job; the context must stay alive while the batch uses it and must
 * be on the batch's device.  *index receives the job's position. */
int pt_batch_add(pt_batch* b, pt_ctx* job, uint32_t* index);
int pt_batch_size(const pt_batch* b, uint32_t* n);

/* One frame of every job: job i renders n_passes[i] passes of params[i]
 * (whole frame, as pt_trace with a 0 x 0 section), with seeds[i] (n_passes[i]
 * entries, or NULL for params[i].seed).  seeds may be NULL.  stats (optional)
 * receives one pt_stats per job; stats[i].kernel_ms is the whole batch's GPU
 * time.  n_passes[i] = 0 skips job i this frame. */
int pt_batch_frame(pt_batch* b, const pt_scene_params* params, const uint32_t* n_passes,
                   const uint32_t* const* seeds, pt_stats* stats);

/* Graph statistics: frames launched, graphs instantiated, in-place updates. */
int pt_batch_graph_info(const pt_batch* b, uint32_t* frames, uint32_t* instantiations, uint32_t* updates);

#ifdef __cplusplus
}
#endif

#endif /* PT_BATCH_H */
