// step.hip -- one iteration of the reference's main loop as one call, run
// eagerly or replayed from a HIP graph.
//
// Reference: devel-izem/src/izem.f90:242-389 (timestep, then with chemistry
// the Strang split reaction_selec(dt/2) -> rk3(dt) -> reaction_selec(dt/2),
// :254-277).  dt stays in the context's time state on the device, so the
// whole iteration is a fixed sequence of kernel launches with no host round
// trip: it is captured once into a graph and replayed, which removes the
// launch gaps between its ~200 kernels.  The graph holds the pointers and
// the chemistry settings it was captured with; izem_step recaptures when
// either changes.
#include "izem_ctx.h"

#include <string.h>

int iz_chem_signature(izem_ctx *ctx, double sig[IZ_CHEM_SIG]);

static int step_body(izem_ctx *c, const double *dx_i, const double *dy_i, const double *dz_i,
                     double *T, double *W_i, double *cp, double *ha, uint8_t *sensor, double *v,
                     int react) {
  int r = izem_timestep(c, dx_i, dy_i, dz_i, cp, W_i, T, v);
  if (r) return r;
  const double *dt = c->ts + 4;  // dtmin of the time state
  if (react && (r = izem_reaction_selec(c, dt, 0.5, T, W_i, cp, ha, v))) return r;
  if ((r = izem_rk3(c, dt, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v))) return r;
  if (react && (r = izem_reaction_selec(c, dt, 0.5, T, W_i, cp, ha, v))) return r;
  return IZEM_OK;
}

void iz_graph_destroy(izem_ctx *c) {
  if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
  if (c->graph.g) (void)hipGraphDestroy(c->graph.g);
  memset(&c->graph, 0, sizeof c->graph);
}

extern "C" int izem_step(izem_ctx *c, const double *dx_i, const double *dy_i, const double *dz_i,
                         double *T, double *W_i, double *cp, double *ha, uint8_t *sensor,
                         double *v, int use_graph) {
  if (!c || !dx_i || !dy_i || !dz_i || !T || !W_i || !cp || !ha || !sensor || !v ||
      use_graph < 0 || use_graph > 1)
    return IZEM_E_ARG;
  const int react = c->o.reaction && c->chem;
  if (!use_graph) return step_body(c, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v, react);
  // what the graph bakes in: the arrays, and the chemistry's kernel variant,
  // launch shape and tolerances (kernel arguments)
  IzGraph key;
  memset(&key, 0, sizeof key);
  const void *ptrs[IZ_GRAPH_PTRS] = {dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v};
  memcpy(key.ptr, ptrs, sizeof ptrs);
  key.react = react;
  key.flux_mode = c->flux_mode;
  if (react) {
    int r = iz_chem_signature(c, key.chem);
    if (r) return r;  // e.g. the stiffness order, whose buffers may grow
  }
  IzGraph &g = c->graph;
  const bool same = g.exec && !memcmp(key.ptr, g.ptr, sizeof key.ptr) && key.react == g.react &&
                    key.flux_mode == g.flux_mode && !memcmp(key.chem, g.chem, sizeof key.chem);
  if (!same) {
    iz_graph_destroy(c);
    // capture on the context's own stream (never the null stream), with the
    // event profiler off; whatever the caller queued on its stream before
    // has finished when the capture starts
    IZ_CHECK(hipStreamSynchronize(c->stream));
    hipStream_t saved = c->stream;
    const int prof = c->prof.on;
    c->stream = c->own_stream;
    c->prof.on = 0;
    hipGraph_t graph = nullptr;
    int r = IZEM_OK;
    if (hipStreamBeginCapture(c->own_stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
      r = IZEM_E_HIP;
    } else {
      r = step_body(c, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v, react);
      if (hipStreamEndCapture(c->own_stream, &graph) != hipSuccess && r == IZEM_OK) r = IZEM_E_HIP;
    }
    c->stream = saved;
    c->prof.on = prof;
    if (r == IZEM_OK && hipGraphInstantiate(&key.exec, graph, nullptr, nullptr, 0) != hipSuccess)
      r = IZEM_E_HIP;
    if (r != IZEM_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return r;
    }
    key.g = graph;
    g = key;
  }
  IZ_CHECK(hipGraphLaunch(g.exec, c->stream));
  return IZEM_OK;
}
