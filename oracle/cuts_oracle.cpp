// CPU restatement of so3g's ML cut handling (src/array_ops.cxx:130-400):
// process_cuts (measure / insert / extract / clear, models "full" and
// "poly") and translate_cuts.  Test infrastructure only: the GPU kernels
// of so3g_amd/csrc/cuts.hip are checked against it.
//
// Ranges arrive flattened: seg[n_seg][3] = (det, start, end) in the order
// the RangesMatrix lists them (detector by detector, ranges ascending), the
// order the reference walks them.  Arithmetic is in T with the reference's
// expression order (array_ops.cxx:312-317, 335-346).
#include <algorithm>
#include <cstdint>
#include <vector>

namespace {

// array_ops.cxx:233-240
int npoly(int gap, int resolution, int nmax) {
    if (nmax < 1) nmax = 1;
    if (gap <= 1) return std::min(1, nmax);
    if (gap <= 3) return std::min(2, nmax);
    if (gap <= 6) return std::min(3, nmax);
    if (gap <= 20) return std::min(4, nmax);
    return std::min(5 + gap / resolution, nmax);
}

template <typename T>
void full_extract(const int32_t *seg, int64_t n_seg, const T *tod, int64_t nsamp, T *vals) {
    int64_t i = 0;
    for (int64_t k = 0; k < n_seg; k++)
        for (int j = seg[3 * k + 1]; j < seg[3 * k + 2]; j++, i++)
            vals[i] = tod[seg[3 * k] * nsamp + j];
}

template <typename T>
void full_insert(const int32_t *seg, int64_t n_seg, T *tod, int64_t nsamp, const T *vals) {
    int64_t i = 0;
    for (int64_t k = 0; k < n_seg; k++)
        for (int j = seg[3 * k + 1]; j < seg[3 * k + 2]; j++, i++)
            tod[seg[3 * k] * nsamp + j] = vals[i];
}

template <typename T>
void poly_extract(const int32_t *seg, int64_t n_seg, int res, int nmax, const T *tod,
                  int64_t nsamp, T *vals) {
    int64_t i = 0;
    for (int64_t k = 0; k < n_seg; k++) {
        const int d = seg[3 * k], a = seg[3 * k + 1], b = seg[3 * k + 2];
        const int np = npoly(b - a, res, nmax);
        if (np <= 1) {
            for (int s = a; s < b; s++) vals[i] += tod[d * nsamp + s];
            i++;
        } else {
            for (int p = 0; p < np; p++) vals[i + p] = 0;
            for (int s = a; s < b; s++) {
                T x = -1 + 2 * (s - a) / T(b - a - 1);
                T t = tod[d * nsamp + s];
                vals[i] += t;
                if (np > 1) vals[i + 1] += t * x;
                if (np > 2) {
                    T Pa = x, Pb = 1, Pc = 0;
                    for (int p = 2; p < np; p++) {
                        Pc = Pb; Pb = Pa; Pa = ((2 * p - 1) * x * Pb - (p - 1) * Pc) / p;
                        vals[i + p] += t * Pa;
                    }
                }
            }
            i += np;
        }
    }
}

template <typename T>
void poly_insert(const int32_t *seg, int64_t n_seg, int res, int nmax, T *tod, int64_t nsamp,
                 const T *vals) {
    int64_t i = 0;
    for (int64_t k = 0; k < n_seg; k++) {
        const int d = seg[3 * k], a = seg[3 * k + 1], b = seg[3 * k + 2];
        const int np = npoly(b - a, res, nmax);
        if (np <= 1) {
            for (int s = a; s < b; s++) tod[d * nsamp + s] = vals[i];
            i++;
        } else {
            for (int s = a; s < b; s++) {
                T x = -1 + 2 * (s - a) / T(b - a - 1);
                T t = vals[i];
                if (np > 1) t += x * vals[i + 1];
                if (np > 2) {
                    T Pa = x, Pb = 1, Pc = 0;
                    for (int p = 2; p < np; p++) {
                        Pc = Pb; Pb = Pa; Pa = ((2 * p - 1) * x * Pb - (p - 1) * Pc) / p;
                        t += Pa * vals[i + p];
                    }
                }
                tod[d * nsamp + s] = t;
            }
            i += np;
        }
    }
}

template <typename T>
void full_translate(const int32_t *iseg, const int32_t *oseg, int64_t n_seg, int64_t insamp,
                    const T *ivals, int64_t onsamp, T *ovals) {
    int64_t ioff = 0, ooff = 0;
    for (int64_t k = 0; k < n_seg; k++) {
        const int64_t i1 = iseg[3 * k + 1], i2 = iseg[3 * k + 2];
        const int64_t o1 = oseg[3 * k + 1], o2 = oseg[3 * k + 2];
        if (onsamp >= insamp) {
            for (int64_t os = o1, oi = ooff; os < o2; os++, oi++) {
                const int64_t is = os * insamp / onsamp;
                ovals[oi] = ivals[ioff + is - i1];
            }
        } else {
            std::vector<int> hits(o2 - o1);
            for (int64_t is = i1, ii = ioff; is < i2; is++, ii++) {
                const int64_t os = is * onsamp / insamp;
                const int64_t oi = ooff + os - o1;
                ovals[oi] += ivals[ii];
                hits[oi - ooff]++;
            }
            for (int64_t os = o1, oi = ooff; os < o2; os++, oi++) ovals[oi] /= hits[oi - ooff];
        }
        ioff += i2 - i1;
        ooff += o2 - o1;
    }
}

template <typename T>
void poly_translate(const int32_t *iseg, const int32_t *oseg, int64_t n_seg, int res, int nmax,
                    const T *ivals, T *ovals) {
    int64_t ioff = 0, ooff = 0;
    for (int64_t k = 0; k < n_seg; k++) {
        const int inp = npoly(iseg[3 * k + 2] - iseg[3 * k + 1], res, nmax);
        const int onp = npoly(oseg[3 * k + 2] - oseg[3 * k + 1], res, nmax);
        int i;
        for (i = 0; i < std::min(inp, onp); i++) ovals[ooff + i] = ivals[ioff + i];
        for (; i < onp; i++) ovals[ooff + i] = 0;
        ioff += inp;
        ooff += onp;
    }
}

template <typename T>
int process(const int32_t *seg, int64_t n_seg, int op, int model, int res, int nmax, T *tod,
            int64_t nsamp, T *vals) {
    if (op == 0) {
        if (model == 0) full_insert(seg, n_seg, tod, nsamp, vals);
        else poly_insert(seg, n_seg, res, nmax, tod, nsamp, vals);
    } else if (op == 1) {
        if (model == 0) full_extract(seg, n_seg, tod, nsamp, vals);
        else poly_extract(seg, n_seg, res, nmax, tod, nsamp, vals);
    } else if (op == 2) {
        for (int64_t k = 0; k < n_seg; k++)
            for (int s = seg[3 * k + 1]; s < seg[3 * k + 2]; s++) tod[seg[3 * k] * nsamp + s] = 0;
    }
    return 0;
}

}  // namespace

extern "C" {

int64_t oracle_cuts_measure(const int32_t *seg, int64_t n_seg, int model, int res, int nmax) {
    int64_t n = 0;
    for (int64_t k = 0; k < n_seg; k++) {
        const int gap = seg[3 * k + 2] - seg[3 * k + 1];
        n += model == 0 ? gap : npoly(gap, res, nmax);
    }
    return n;
}

int oracle_process_cuts(const int32_t *seg, int64_t n_seg, int op, int model, int res, int nmax,
                        void *tod, int64_t nsamp, void *vals, int f64) {
    if (f64) return process(seg, n_seg, op, model, res, nmax, (double *)tod, nsamp, (double *)vals);
    return process(seg, n_seg, op, model, res, nmax, (float *)tod, nsamp, (float *)vals);
}

int oracle_translate_cuts(const int32_t *iseg, const int32_t *oseg, int64_t n_seg, int model,
                          int res, int nmax, int64_t insamp, int64_t onsamp, const void *ivals,
                          void *ovals, int f64) {
    if (model == 0) {
        if (f64) full_translate(iseg, oseg, n_seg, insamp, (const double *)ivals, onsamp, (double *)ovals);
        else full_translate(iseg, oseg, n_seg, insamp, (const float *)ivals, onsamp, (float *)ovals);
    } else {
        if (f64) poly_translate(iseg, oseg, n_seg, res, nmax, (const double *)ivals, (double *)ovals);
        else poly_translate(iseg, oseg, n_seg, res, nmax, (const float *)ivals, (float *)ovals);
    }
    return 0;
}

}  // extern "C"
