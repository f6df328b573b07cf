/*
 * dlp_tune.h -- the process-wide tuning table (internal; dlp_autotune.hip).
 */
#ifndef DLP_TUNE_H
#define DLP_TUNE_H

#include "dynlp/dynlp_c.h"

/* loads <library directory>/tuning/<gfx arch>.tune once per process (no
 * effect when the file is absent) */
void dlp_tuning_table_autoload();
/* the table's tuning for a context of this shape, if any */
bool dlp_tuning_lookup_ctx(const dlp_shape_key &key, dlp_tuning *out);
/* the batched runner's shape: log2 of the graph count and of the total
 * vertex count (rounded up) */
void dlp_multi_shape(int numGraphs, long long totalNodes, int *logGraphs, int *logNodes);
bool dlp_tuning_lookup_multi(int logGraphs, int logNodes, dlp_multi_tuning *out);
void dlp_tuning_put_multi(int logGraphs, int logNodes, const dlp_multi_tuning &t, float ms);

#endif /* DLP_TUNE_H */
