/* host_internal.h -- shared state of the host-side helpers (not exported). */
#ifndef DLP_HOST_INTERNAL_H
#define DLP_HOST_INTERNAL_H

namespace dlp_host {
/* Suppresses the generators' progress printing (the C-ABI input generator
 * sets it; the CLI keeps the reference's output). */
extern bool g_quiet;
}  // namespace dlp_host

#endif
