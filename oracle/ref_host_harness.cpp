// Test-only C entry points over the reference's own host sources, compiled from
// /root/reference/src by oracle/Makefile (target ref): TimeSeries::fromCsv / interp
// (src/timeseries.cpp) and utils::solveSimpleMathExpression / miniFormat (src/utils.h).
// Used by tests/ to pin the fip CLI's parsing against the reference implementation.
#include <cstring>
#include <stdexcept>
#include <string>

#include "timeseries.h"
#include "utils.h"

extern "C" {

// Loads a CSV series and evaluates it at n times. Returns 0, or -1 with the exception
// text in err.
int ref_series_interp(const char* path, int tcol, int dcol, const float* t, int n, float* out,
                      char* err, int errcap) {
  try {
    TimeSeries* ts = TimeSeries::fromCsv(path, tcol, dcol);
    for (int i = 0; i < n; i++) out[i] = ts->interp(t[i]);
    delete ts;
    return 0;
  } catch (std::exception& e) {
    std::strncpy(err, e.what(), errcap - 1);
    err[errcap - 1] = 0;
    return -1;
  }
}

int ref_math_expr(const char* expr, double* out, char* err, int errcap) {
  try {
    *out = utils::solveSimpleMathExpression(expr);
    return 0;
  } catch (std::exception& e) {
    std::strncpy(err, e.what(), errcap - 1);
    err[errcap - 1] = 0;
    return -1;
  }
}
}
