// flow_only.hip -- the two hooks flow.hip calls into the chemistry and
// graph translation units, for the flow-only library variants (other
// species counts or WENO7; phd_amd/build.py): a flow-only context never
// holds chemistry or a captured step.
#include "izem_ctx.h"

void iz_chem_destroy(izem_ctx *ctx) { (void)ctx; }
void iz_graph_destroy(izem_ctx *ctx) { (void)ctx; }
