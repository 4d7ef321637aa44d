// Column-panel layout of a single instance's CSR (device side, at handle
// creation). Panel p of P covers columns [p n / P, (p + 1) n / P); the panels
// are stored one after another, each a CSR over all n rows whose entries keep
// the row's stored order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hgqp {

// Bytes of the gathered vector per panel: half of one XCD's 4 MiB L2, the
// fastest panel size measured for the FP32 and FP64 passes over a matrix
// whose columns are spread over all n.
constexpr int64_t kPanelBytes = int64_t(2) << 20;

// An entry is far when its column is more than kPanelReach from its row;
// panels only pay when the matrix gathers far from the diagonal.
constexpr int64_t kPanelReach = 65536;

struct PanelSurvey {
    int64_t far = 0;     // far entries
    bool sorted = true;  // every row's columns are nondecreasing
};

// Far entries and column order of a device CSR.
PanelSurvey survey_panels(int64_t n, const int* start, const int* col, hipStream_t stream);

// Panel count for a vector of `elem`-byte values: n elem / kPanelBytes
// rounded, in [1, kMaxPanels], and 1 unless the survey shows sorted rows with
// at least 8 far entries per row on average. The panel count changes no FP64
// ordered result: the passes add each row's products in stored order.
int panels_for(int64_t n, const PanelSurvey& s, int elem);

// Splits the device CSR (start, col, val) into `count` panels: writes
// start_out[count (n + 1)] (offsets into col_out / val_out), col_out[nnz] and
// val_out[nnz] (rounded to T). Requires sorted rows and count n < 2^31.
template <class T>
void split_panels(int64_t n, int64_t nnz, int count, const int* start, const int* col, const double* val,
                  int* start_out, int* col_out, T* val_out, hipStream_t stream);

}  // namespace hgqp
