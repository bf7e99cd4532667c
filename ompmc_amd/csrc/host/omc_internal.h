/* Private helpers shared by the host translation units. */
#ifndef OMC_INTERNAL_H
#define OMC_INTERNAL_H

#include <stdarg.h>

int omc_set_error(int code, const char *fmt, ...);

#define OMC_TRY(expr) do { int rc_ = (expr); if (rc_ < 0) return rc_; } while (0)

#endif
