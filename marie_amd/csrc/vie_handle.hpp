// Operator handle shared by the matvec, assembly and Krylov translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "vie_fft.hpp"

namespace marie {

enum Status {
    OK = 0,
    ERR_ARG = 1,
    ERR_ALLOC = 2,
    ERR_HIP = 3,
    ERR_STATE = 4,
    ERR_SIZE = 5,
    ERR_NONFINITE = 6,
};

enum Precision { PREC_F64 = 1, PREC_F32 = 2 };
// OP_P: the JVIE preconditioner, mv_AN with the material (1/Mr, 1/Mr - 1)
enum OpKind { OP_AN = 0, OP_N = 1, OP_L = 2, OP_K = 3, OP_P = 4 };

void set_error(const std::string& msg);
const char* get_error();

#define MARIE_HIP(call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::marie::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
            return ::marie::ERR_HIP;                                                       \
        }                                                                                  \
    } while (0)

struct Op {
    int L = 0, M = 0, N = 0;        // voxel grid
    int Px = 0, Py = 0, Pz = 0;     // circulant (FFT) grid
    int precisions = 0;             // bitmask of Precision
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    AxisPlan plan[3];

    // device data
    // fN: full c2<T>[6][Pz][Py][Px], or its octant [Pz/2+1][Py/2+1][Px/2+1][6]
    // (the 6 components of a point adjacent, so the z pass reads one 96-byte
    // run per point) when fN_oct (every component has the circulant's per-axis
    // parity)
    void* fN64 = nullptr;
    void* fN32 = nullptr;
    bool fN_oct = false;
    // fK: the K operator, c2<T>[3][Pz][Py][Px] or its octant [..][3] (fK_oct)
    void* fK64 = nullptr;
    void* fK32 = nullptr;
    bool fK_oct = false;
    bool have_fK = false;
    void* tw64 = nullptr;           // 3 axis tables, offsets 0, Px, Px+Py
    void* tw32 = nullptr;
    void* Mr64 = nullptr;           // c2<double>[N][M][L]
    void* Mc64 = nullptr;
    void* Mr32 = nullptr;
    void* Mc32 = nullptr;
    double Gram = 0.0;
    bool have_fN = false, have_mat = false, have_idx = false;
    // preconditioner material 1/Mr, 1/Mr - 1 (derived from Mr on first use);
    // mr_unit: every Mr is exactly 1 (form II), so the preconditioner is Gram*x
    void* Pr64 = nullptr;
    void* Pc64 = nullptr;
    void* Pr32 = nullptr;
    void* Pc32 = nullptr;
    bool have_P = false, mr_unit = false;

    int* pos = nullptr;             // [3 L M N] -> compact position or -1
    int64_t* idx = nullptr;         // [nidx] 0-based grid index of each compact entry
    int64_t nidx = 0;

    // work buffers (bytes), sized for `nrhs_cap` right-hand sides at the widest precision
    void* A = nullptr;              // [3 nrhs][N][M][Px]
    void* B = nullptr;              // [3 nrhs][N][Py][Px]
    int nrhs_cap = 0;

    // per-kernel variants (fwd_x, fwd_y, z, inv_y, inv_x): the register engine
    // where the side allows it (apply falls back to the LDS-stage engine)
    // right-hand sides per x/y launch of a multi-column product (0: as many
    // as keep their x-pass buffers within kXYGroupBytes; MARIE_XY_GROUP at
    // creation sets it)
    int xy_group = 0;
    int emax64[5] = {-4, -16, -8, -208, -4};
    int emax32[5] = {-8, -16, -16, -16, -8};

    // optional per-kernel timing with HIP events on the handle's stream: each
    // product records 6 events into the next row of a ring, read back (with
    // one synchronisation) when the ring is full or the timings are asked
    // for, so timed products still run back to back
    bool timing = false;
    hipEvent_t ev[8] = {};
    static constexpr int kTimingRing = 64;
    hipEvent_t ring[kTimingRing][6] = {};
    int ring_used = 0;
    double kernel_ms[8] = {};
    int64_t kernel_calls = 0;

    int64_t nvox() const { return (int64_t)L * M * N; }
};

// bytes of x-pass output the multi-column product keeps per group of
// columns: below the 256 MB last-level cache, so the y pass reads it there
constexpr int64_t kXYGroupBytes = (int64_t)192 << 20;

int ensure_work(Op* h, int nrhs);
int apply(Op* h, int op, int transp, int prec, const void* x, void* y, int nrhs, hipStream_t s);
// oscale * (the product) for MARIE_OP_AN, N, L, K: the factor enters the
// epilogue's coefficients (used to keep a Krylov power basis near unit norm)
int apply_scaled(Op* h, int op, int transp, int prec, const void* x, void* y, int nrhs, hipStream_t s,
                 double oscale);
// allocation-free product on the per-device scratch array (vie_mvp.hip)
int scratch_init(int device);
size_t scratch_capacity();
size_t noalloc_bytes(const Op* h, int prec, int nrhs);
int apply_noalloc(Op* h, int op, int transp, int prec, const void* x, void* y, int nrhs, hipStream_t s);
int build_plan(int P, AxisPlan* pl);
int upload_twiddles(Op* h);
int store_fN(Op* h, void* full64);   // takes ownership of a device complex128 fN
int load_fN(Op* h, void* out64);     // full complex128 fN into a device buffer
int store_circ(Op* h, int kind, void* full64);   // kind 0: fN, 1: fK
int load_circ(Op* h, int kind, void* out64);
int assemble_K(Op* h, double dx, double ko, int stage = -1, double* out = nullptr);
int autotune(Op* h, int prec, int iters, double* best_ms5);
int timing_flush(Op* h);   // folds the recorded ring rows into kernel_ms
// refinement with s-step GMRES cycles on complex64 products (vie_sstep.hip)
int sstep_refine(Op* h, hipStream_t s, int transp, const c2<double>* b, c2<double>* x, c2<double>* r, c2<double>* t,
                 double bnorm, double rnorm, int m, double tol, int maxit, double inner_tol,
                 std::vector<double>& resvec, int* p_out, int* outit_out);

}  // namespace marie
