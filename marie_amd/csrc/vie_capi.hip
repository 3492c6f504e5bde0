// extern "C" entry points of libmarie_vie.so (declared in include/marie_vie.h).
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

struct marie_vie_op : marie::Op {};

namespace marie {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

template <typename T> __global__ void k_d2f(const c2<double>* __restrict__ s, c2<T>* __restrict__ d, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = c2<T>{(T)s[i].x, (T)s[i].y};
}

static int to_device_both(Op* h, const void* src_host64, size_t n, void** d64, void** d32, bool from_device) {
    // uploads complex128 data and derives the complex64 copy on the device
    if (h->precisions & PREC_F64) {
        if (!*d64 && hipMalloc(d64, n * sizeof(c2<double>)) != hipSuccess) {
            (void)hipGetLastError();
            *d64 = nullptr;
            set_error("out of device memory");
            return ERR_ALLOC;
        }
    }
    void* stage = *d64;
    bool temp = false;
    if (!stage) {
        MARIE_HIP(hipMalloc(&stage, n * sizeof(c2<double>)));
        temp = true;
    }
    MARIE_HIP(hipMemcpyAsync(stage, src_host64, n * sizeof(c2<double>),
                             from_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->stream));
    if (h->precisions & PREC_F32) {
        if (!*d32 && hipMalloc(d32, n * sizeof(c2<float>)) != hipSuccess) {
            (void)hipGetLastError();
            *d32 = nullptr;
            set_error("out of device memory");
            return ERR_ALLOC;
        }
        hipLaunchKernelGGL(k_d2f<float>, dim3(2048), dim3(256), 0, h->stream, (const c2<double>*)stage,
                           (c2<float>*)*d32, (int64_t)n);
        MARIE_HIP(hipGetLastError());
    }
    MARIE_HIP(hipStreamSynchronize(h->stream));
    if (temp) MARIE_HIP(hipFree(stage));
    return OK;
}

}  // namespace marie

using namespace marie;

#define CHECK_H(h)                                   \
    do {                                             \
        if (!(h)) {                                  \
            set_error("null handle");                \
            return MARIE_ERR_ARG;                    \
        }                                            \
        MARIE_HIP(hipSetDevice((h)->device));        \
    } while (0)

extern "C" {

int marie_vie_version(void) { return 10000; }

const char* marie_vie_last_error(void) { return get_error(); }

int marie_vie_create(marie_vie_op** out, int L, int M, int N, int Px, int Py, int Pz, int precisions,
                     int device) {
    if (!out) { set_error("null output pointer"); return MARIE_ERR_ARG; }
    *out = nullptr;
    if (L < 1 || M < 1 || N < 1 || Px < L || Py < M || Pz < N) {
        set_error("invalid grid / circulant sizes (need 1 <= L <= Px, ...)");
        return MARIE_ERR_ARG;
    }
    if (precisions <= 0 || (precisions & ~(PREC_F64 | PREC_F32))) {
        set_error("precisions must be a non-empty mask of MARIE_PREC_F64|MARIE_PREC_F32");
        return MARIE_ERR_ARG;
    }
    if ((int64_t)3 * L * M * N >= (int64_t)1 << 31) {
        set_error("grid too large for 32-bit compact positions");
        return MARIE_ERR_SIZE;
    }
    int ndev = 0;
    MARIE_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) { set_error("invalid device ordinal"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(device));
    auto* h = new marie_vie_op();
    h->L = L; h->M = M; h->N = N; h->Px = Px; h->Py = Py; h->Pz = Pz;
    h->precisions = precisions;
    h->device = device;
    if (const char* g = std::getenv("MARIE_XY_GROUP")) h->xy_group = std::atoi(g);
    int st = OK;
    for (int a = 0; a < 3 && st == OK; ++a) st = build_plan(a == 0 ? Px : a == 1 ? Py : Pz, &h->plan[a]);
    if (st == OK && hipStreamCreate(&h->stream) != hipSuccess) {
        set_error("hipStreamCreate failed");
        st = MARIE_ERR_HIP;
    }
    if (st == OK) h->own_stream = true;
    if (st == OK) st = scratch_init(device);     // the allocation-free product's per-device array
    if (st == OK) st = upload_twiddles(h);
    for (int k = 0; k < 8 && st == OK; ++k)
        if (hipEventCreate(&h->ev[k]) != hipSuccess) { set_error("hipEventCreate failed"); st = MARIE_ERR_HIP; }
    if (st != OK) {
        marie_vie_destroy(h);
        return st;
    }
    *out = h;
    return MARIE_OK;
}

int marie_vie_destroy(marie_vie_op* h) {
    if (!h) return MARIE_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* bufs[] = {h->fN64, h->fN32, h->fK64, h->fK32, h->tw64, h->tw32, h->Mr64, h->Mc64, h->Mr32, h->Mc32,
                    h->Pr64, h->Pc64, h->Pr32, h->Pc32, h->pos, h->idx, h->A, h->B};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int k = 0; k < 8; ++k)
        if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
    for (auto& row : h->ring)
        for (hipEvent_t e : row)
            if (e) (void)hipEventDestroy(e);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return MARIE_OK;
}

int marie_vie_set_stream(marie_vie_op* h, void* stream) {
    CHECK_H(h);
    MARIE_HIP(hipStreamSynchronize(h->stream));
    if (stream) {
        if (h->own_stream) MARIE_HIP(hipStreamDestroy(h->stream));
        h->own_stream = false;
        h->stream = (hipStream_t)stream;
    } else if (!h->own_stream) {
        MARIE_HIP(hipStreamCreate(&h->stream));
        h->own_stream = true;
    }
    return MARIE_OK;
}

int marie_vie_set_fN(marie_vie_op* h, const double* fN) {
    CHECK_H(h);
    if (!fN) { set_error("null fN"); return MARIE_ERR_ARG; }
    const size_t n = (size_t)6 * h->Px * h->Py * h->Pz;
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(fN[i])) { set_error("fN has non-finite entries"); return MARIE_ERR_NONFINITE; }
    void* full = nullptr;
    if (hipMalloc(&full, n * sizeof(c2<double>)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("out of device memory for fN");
        return MARIE_ERR_ALLOC;
    }
    if (hipMemcpy(full, fN, n * sizeof(c2<double>), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(full);
        set_error("fN upload failed");
        return MARIE_ERR_HIP;
    }
    return store_fN(h, full);
}

int marie_vie_set_fN_device(marie_vie_op* h, const void* fN_dev) {
    CHECK_H(h);
    if (!fN_dev) { set_error("null fN"); return MARIE_ERR_ARG; }
    const size_t n = (size_t)6 * h->Px * h->Py * h->Pz;
    void* full = nullptr;
    if (hipMalloc(&full, n * sizeof(c2<double>)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("out of device memory for fN");
        return MARIE_ERR_ALLOC;
    }
    if (hipMemcpyAsync(full, fN_dev, n * sizeof(c2<double>), hipMemcpyDeviceToDevice, h->stream) != hipSuccess) {
        (void)hipFree(full);
        set_error("fN copy failed");
        return MARIE_ERR_HIP;
    }
    return store_fN(h, full);
}

int marie_vie_get_fN(marie_vie_op* h, double* fN) {
    CHECK_H(h);
    if (!h->have_fN || !h->fN64) { set_error("no FP64 fN on the handle"); return MARIE_ERR_STATE; }
    const size_t n = (size_t)6 * h->Px * h->Py * h->Pz;
    void* full = nullptr;
    MARIE_HIP(hipMalloc(&full, n * sizeof(c2<double>)));
    int st = load_fN(h, full);
    if (st == OK && hipMemcpyAsync(fN, full, n * sizeof(c2<double>), hipMemcpyDeviceToHost, h->stream) != hipSuccess) {
        set_error("fN download failed");
        st = MARIE_ERR_HIP;
    }
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(full);
    return st;
}

int marie_vie_set_fK(marie_vie_op* h, const double* fK) {
    CHECK_H(h);
    if (!fK) { set_error("null fK"); return MARIE_ERR_ARG; }
    const size_t n = (size_t)3 * h->Px * h->Py * h->Pz;
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(fK[i])) { set_error("fK has non-finite entries"); return MARIE_ERR_NONFINITE; }
    void* full = nullptr;
    if (hipMalloc(&full, n * sizeof(c2<double>)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("out of device memory for fK");
        return MARIE_ERR_ALLOC;
    }
    if (hipMemcpy(full, fK, n * sizeof(c2<double>), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(full);
        set_error("fK upload failed");
        return MARIE_ERR_HIP;
    }
    return store_circ(h, 1, full);
}

int marie_vie_get_fK(marie_vie_op* h, double* fK) {
    CHECK_H(h);
    if (!h->have_fK || !h->fK64) { set_error("no FP64 fK on the handle"); return MARIE_ERR_STATE; }
    const size_t n = (size_t)3 * h->Px * h->Py * h->Pz;
    void* full = nullptr;
    MARIE_HIP(hipMalloc(&full, n * sizeof(c2<double>)));
    int st = load_circ(h, 1, full);
    if (st == OK && hipMemcpyAsync(fK, full, n * sizeof(c2<double>), hipMemcpyDeviceToHost, h->stream) != hipSuccess) {
        set_error("fK download failed");
        st = MARIE_ERR_HIP;
    }
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(full);
    return st;
}

int marie_vie_assemble_K(marie_vie_op* h, double dx, double ko) {
    CHECK_H(h);
    return assemble_K(h, dx, ko);
}

int marie_vie_set_material(marie_vie_op* h, const double* Mr, const double* Mc, double Gram) {
    CHECK_H(h);
    if (!Mr || !Mc) { set_error("null material"); return MARIE_ERR_ARG; }
    const size_t n = (size_t)h->nvox();
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(Mr[i]) || !std::isfinite(Mc[i])) {
            set_error("material has non-finite entries");
            return MARIE_ERR_NONFINITE;
        }
    if (!std::isfinite(Gram)) { set_error("non-finite Gram"); return MARIE_ERR_NONFINITE; }
    int st = to_device_both(h, Mr, n, &h->Mr64, &h->Mr32, false);
    if (st) return st;
    st = to_device_both(h, Mc, n, &h->Mc64, &h->Mc32, false);
    if (st) return st;
    h->Gram = Gram;
    h->have_mat = true;
    h->have_P = false;
    h->mr_unit = true;
    for (size_t i = 0; i < n && h->mr_unit; ++i) h->mr_unit = Mr[2 * i] == 1.0 && Mr[2 * i + 1] == 0.0;
    return MARIE_OK;
}

int marie_vie_set_index(marie_vie_op* h, const int64_t* idx, int64_t nidx, int one_based) {
    CHECK_H(h);
    const int64_t ng = 3 * h->nvox();
    if (nidx < 0 || nidx > ng || (nidx > 0 && !idx)) { set_error("invalid index length"); return MARIE_ERR_ARG; }
    std::vector<int> pos((size_t)ng, -1);
    std::vector<int64_t> z((size_t)nidx);
    const int64_t base = one_based ? 1 : 0;
    for (int64_t t = 0; t < nidx; ++t) {
        const int64_t g = idx[t] - base;
        if (g < 0 || g >= ng) { set_error("index out of range"); return MARIE_ERR_ARG; }
        if (pos[(size_t)g] >= 0) { set_error("duplicate index entry"); return MARIE_ERR_ARG; }
        pos[(size_t)g] = (int)t;
        z[(size_t)t] = g;
    }
    if (!h->pos) MARIE_HIP(hipMalloc(&h->pos, (size_t)ng * sizeof(int)));
    if (h->idx) { MARIE_HIP(hipFree(h->idx)); h->idx = nullptr; }
    if (nidx) MARIE_HIP(hipMalloc(&h->idx, (size_t)nidx * sizeof(int64_t)));
    MARIE_HIP(hipMemcpyAsync(h->pos, pos.data(), (size_t)ng * sizeof(int), hipMemcpyHostToDevice, h->stream));
    if (nidx)
        MARIE_HIP(hipMemcpyAsync(h->idx, z.data(), (size_t)nidx * sizeof(int64_t), hipMemcpyHostToDevice,
                                 h->stream));
    MARIE_HIP(hipStreamSynchronize(h->stream));
    h->nidx = nidx;
    h->have_idx = true;
    return MARIE_OK;
}

int marie_vie_apply_device(marie_vie_op* h, int op, int transp, int prec, const void* x, void* y, int nrhs) {
    CHECK_H(h);
    if (op < MARIE_OP_AN || op > MARIE_OP_P || (prec != MARIE_PREC_F64 && prec != MARIE_PREC_F32) || nrhs < 1 ||
        (h->nidx > 0 && (!x || !y))) {
        set_error("invalid apply arguments");
        return MARIE_ERR_ARG;
    }
    if (!(op == MARIE_OP_K ? h->have_fK : h->have_fN) || !h->have_idx ||
        ((op == MARIE_OP_AN || op == MARIE_OP_P) && !h->have_mat)) {
        set_error("operator, index or material not set");
        return MARIE_ERR_STATE;
    }
    if (h->nidx == 0) return MARIE_OK;
    return apply(h, op, transp ? 1 : 0, prec, x, y, nrhs, h->stream);
}

int marie_vie_apply_device_noalloc(marie_vie_op* h, int op, int transp, int prec, const void* x, void* y,
                                   int nrhs) {
    CHECK_H(h);
    if (op < MARIE_OP_AN || op > MARIE_OP_P || (prec != MARIE_PREC_F64 && prec != MARIE_PREC_F32) || nrhs < 1 ||
        (h->nidx > 0 && (!x || !y))) {
        set_error("invalid apply arguments");
        return MARIE_ERR_ARG;
    }
    if (noalloc_bytes(h, prec, nrhs) > scratch_capacity()) {
        set_error("the product's work buffers (" + std::to_string(noalloc_bytes(h, prec, nrhs)) +
                  " bytes) exceed the allocation-free scratch of " + std::to_string(scratch_capacity()) + " bytes");
        return MARIE_ERR_SIZE;
    }
    if (!(op == MARIE_OP_K ? h->have_fK : h->have_fN) || !h->have_idx ||
        ((op == MARIE_OP_AN || op == MARIE_OP_P) && !h->have_mat)) {
        set_error("operator, index or material not set");
        return MARIE_ERR_STATE;
    }
    if (h->nidx == 0) return MARIE_OK;
    return apply_noalloc(h, op, transp ? 1 : 0, prec, x, y, nrhs, h->stream);
}

size_t marie_vie_noalloc_scratch_bytes(void) { return scratch_capacity(); }

int marie_vie_apply(marie_vie_op* h, int op, int transp, const double* x, double* y, int nrhs) {
    CHECK_H(h);
    if (nrhs < 1 || op < MARIE_OP_AN || op > MARIE_OP_P) { set_error("invalid apply arguments"); return MARIE_ERR_ARG; }
    if (!(op == MARIE_OP_K ? h->have_fK : h->have_fN) || !h->have_idx ||
        ((op == MARIE_OP_AN || op == MARIE_OP_P) && !h->have_mat)) {
        set_error("operator, index or material not set");
        return MARIE_ERR_STATE;
    }
    if (h->nidx == 0) return MARIE_OK;
    if (!x || !y) { set_error("null vector"); return MARIE_ERR_ARG; }
    const size_t n = (size_t)h->nidx * nrhs;
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(x[i])) { set_error("input vector has non-finite entries"); return MARIE_ERR_NONFINITE; }
    const bool f64 = h->precisions & PREC_F64;
    void *dx = nullptr, *dy = nullptr;
    MARIE_HIP(hipMalloc(&dx, n * sizeof(c2<double>)));
    if (hipMalloc(&dy, n * sizeof(c2<double>)) != hipSuccess) {
        (void)hipFree(dx);
        set_error("out of device memory");
        return MARIE_ERR_ALLOC;
    }
    int st = OK;
    if (hipMemcpyAsync(dx, x, n * sizeof(c2<double>), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
        set_error("H2D copy failed");
        st = MARIE_ERR_HIP;
    }
    if (st == OK && !f64) {
        // FP32-only handle: convert in place (first half of the buffer)
        hipLaunchKernelGGL(k_d2f<float>, dim3(1024), dim3(256), 0, h->stream, (const c2<double>*)dx,
                           (c2<float>*)dy, (int64_t)n);
        std::swap(dx, dy);
    }
    if (st == OK) st = apply(h, op, transp ? 1 : 0, f64 ? PREC_F64 : PREC_F32, dx, dy, nrhs, h->stream);
    if (st == OK) {
        if (f64) {
            if (hipMemcpyAsync(y, dy, n * sizeof(c2<double>), hipMemcpyDeviceToHost, h->stream) != hipSuccess) {
                set_error("D2H copy failed");
                st = MARIE_ERR_HIP;
            }
        } else {
            std::vector<float> tmp(2 * n);
            if (hipMemcpyAsync(tmp.data(), dy, n * sizeof(c2<float>), hipMemcpyDeviceToHost, h->stream) !=
                hipSuccess) {
                set_error("D2H copy failed");
                st = MARIE_ERR_HIP;
            }
            (void)hipStreamSynchronize(h->stream);
            for (size_t i = 0; i < 2 * n; ++i) y[i] = tmp[i];
        }
    }
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(dx);
    (void)hipFree(dy);
    return st;
}

int marie_vie_enable_timing(marie_vie_op* h, int on) {
    CHECK_H(h);
    MARIE_HIP(hipStreamSynchronize(h->stream));   // recorded rows are dropped
    h->ring_used = 0;
    h->timing = on != 0;
    for (int k = 0; k < 8; ++k) h->kernel_ms[k] = 0.0;
    h->kernel_calls = 0;
    return MARIE_OK;
}

int marie_vie_get_timing(marie_vie_op* h, double* mean_ms5, int64_t* calls) {
    CHECK_H(h);
    const int st = timing_flush(h);
    if (st) return st;
    for (int k = 0; k < 5; ++k) mean_ms5[k] = h->kernel_calls ? h->kernel_ms[k] / h->kernel_calls : 0.0;
    if (calls) *calls = h->kernel_calls;
    return MARIE_OK;
}

}  // extern "C"

extern "C" int marie_vie_autotune(marie_vie_op* h, int prec, int iters, int* cfg6, double* best_ms5) {
    CHECK_H(h);
    if ((prec != MARIE_PREC_F64 && prec != MARIE_PREC_F32) || !(h->precisions & prec) || iters < 1) {
        set_error("invalid autotune arguments");
        return MARIE_ERR_ARG;
    }
    if (!h->have_fN || !h->have_idx || !h->have_mat) {
        set_error("operator, index or material not set");
        return MARIE_ERR_STATE;
    }
    int st = autotune(h, prec, iters, best_ms5);
    if (st) return st;
    if (cfg6) {
        const int* em = prec == MARIE_PREC_F64 ? h->emax64 : h->emax32;
        for (int k = 0; k < 5; ++k) cfg6[k] = em[k];
        cfg6[5] = 0;
    }
    return MARIE_OK;
}

extern "C" int marie_vie_set_tiles(marie_vie_op* h, int prec, const int* cfg6) {
    CHECK_H(h);
    if ((prec != MARIE_PREC_F64 && prec != MARIE_PREC_F32) || !cfg6) {
        set_error("invalid tile arguments");
        return MARIE_ERR_ARG;
    }
    // accepted codes: the LDS-stage engine's elements per thread (positive;
    // a value between compiled variants runs the next larger one), the
    // register engine's points per thread (negative), for the z pass its
    // one-group-per-component codes -(100 * TX + E) and for the y passes the
    // two-lines-per-thread codes -(200 + E)
    static const int kReg[] = {4, 8, 16, 10, 12, 14, 20, 24, 28};
    static const int kZ3[] = {404, 408, 416, 804, 808, 816, 414};
    static const int kPair[] = {204, 208, 216};
    for (int k = 0; k < 5; ++k) {
        const int v = cfg6[k];
        bool ok = v == 4 || v == 8 || v == 16 || v == 32;
        for (int r : kReg) ok = ok || v == -r;
        if (k == 2)
            for (int r : kZ3) ok = ok || v == -r;
        if (k == 1 || k == 3)
            for (int r : kPair) ok = ok || v == -r;
        if (!ok) {
            set_error("tile code " + std::to_string(v) + " of pass " + std::to_string(k) + " is not a compiled variant");
            return MARIE_ERR_ARG;
        }
    }
    int* em = prec == MARIE_PREC_F64 ? h->emax64 : h->emax32;
    for (int k = 0; k < 5; ++k) em[k] = cfg6[k];
    return MARIE_OK;
}
