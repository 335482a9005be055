// Minimal CSR sparse matrix (oracle side; see csr.h).
#include "csr.h"
#include <algorithm>
#include <numeric>

Csr csr_from_triplets(const Triplets& t, int64_t m, int64_t n) {
    Csr a;
    a.nrows = m;
    a.ncols = n;
    const size_t nt = t.i.size();
    std::vector<int64_t> cnt(m + 1, 0);
    for (size_t k = 0; k < nt; ++k) cnt[t.i[k] + 1]++;
    for (int64_t r = 0; r < m; ++r) cnt[r + 1] += cnt[r];
    std::vector<int64_t> cols(nt);
    std::vector<double> vals(nt);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < nt; ++k) {
        int64_t p = pos[t.i[k]]++;
        cols[p] = t.j[k];
        vals[p] = t.v[k];
    }
    a.rowptr.assign(m + 1, 0);
    std::vector<int64_t> idx;
    for (int64_t r = 0; r < m; ++r) {
        int64_t b = cnt[r], e = cnt[r + 1];
        idx.resize(e - b);
        std::iota(idx.begin(), idx.end(), b);
        // stable by column so duplicates are summed in insertion order
        std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return cols[x] < cols[y]; });
        for (size_t q = 0; q < idx.size(); ++q) {
            int64_t c = cols[idx[q]];
            if (!a.col.empty() && (int64_t)a.col.size() > a.rowptr[r] && a.col.back() == c)
                a.val.back() += vals[idx[q]];
            else {
                a.col.push_back(c);
                a.val.push_back(vals[idx[q]]);
            }
        }
        a.rowptr[r + 1] = (int64_t)a.col.size();
    }
    return a;
}

Csr csr_identity(int64_t n) {
    Csr a;
    a.nrows = a.ncols = n;
    a.rowptr.resize(n + 1);
    a.col.resize(n);
    a.val.assign(n, 1.0);
    for (int64_t r = 0; r <= n; ++r) a.rowptr[r] = r;
    for (int64_t r = 0; r < n; ++r) a.col[r] = r;
    return a;
}

Csr csr_add(const Csr& a, const Csr& b, double alpha, double beta) {
    Csr c;
    c.nrows = a.nrows;
    c.ncols = a.ncols;
    c.rowptr.assign(a.nrows + 1, 0);
    for (int64_t r = 0; r < a.nrows; ++r) {
        int64_t p = a.rowptr[r], pe = a.rowptr[r + 1], q = b.rowptr[r], qe = b.rowptr[r + 1];
        while (p < pe || q < qe) {
            if (q >= qe || (p < pe && a.col[p] < b.col[q])) {
                c.col.push_back(a.col[p]);
                c.val.push_back(alpha * a.val[p]);
                ++p;
            } else if (p >= pe || b.col[q] < a.col[p]) {
                c.col.push_back(b.col[q]);
                c.val.push_back(beta * b.val[q]);
                ++q;
            } else {
                c.col.push_back(a.col[p]);
                c.val.push_back(alpha * a.val[p] + beta * b.val[q]);
                ++p;
                ++q;
            }
        }
        c.rowptr[r + 1] = (int64_t)c.col.size();
    }
    return c;
}

Csr csr_mul(const Csr& a, const Csr& b) {
    Csr c;
    c.nrows = a.nrows;
    c.ncols = b.ncols;
    c.rowptr.assign(a.nrows + 1, 0);
    std::vector<int64_t> marker(b.ncols, -1);
    std::vector<double> acc(b.ncols, 0.0);
    std::vector<int64_t> touched;
    for (int64_t r = 0; r < a.nrows; ++r) {
        touched.clear();
        for (int64_t p = a.rowptr[r]; p < a.rowptr[r + 1]; ++p) {
            int64_t k = a.col[p];
            double av = a.val[p];
            for (int64_t q = b.rowptr[k]; q < b.rowptr[k + 1]; ++q) {
                int64_t j = b.col[q];
                if (marker[j] != r) {
                    marker[j] = r;
                    acc[j] = 0.0;
                    touched.push_back(j);
                }
                acc[j] += av * b.val[q];
            }
        }
        std::sort(touched.begin(), touched.end());
        for (int64_t j : touched) {
            c.col.push_back(j);
            c.val.push_back(acc[j]);
        }
        c.rowptr[r + 1] = (int64_t)c.col.size();
    }
    return c;
}

Csr csr_scale(const Csr& a, double s) {
    Csr c = a;
    for (double& v : c.val) v *= s;
    return c;
}

Csr csr_transpose(const Csr& a) {
    Triplets t;
    for (int64_t r = 0; r < a.nrows; ++r)
        for (int64_t p = a.rowptr[r]; p < a.rowptr[r + 1]; ++p) t.push(a.col[p], r, a.val[p]);
    return csr_from_triplets(t, a.ncols, a.nrows);
}

void csr_dropzeros(Csr& a) {
    int64_t w = 0;
    std::vector<int64_t> rp(a.nrows + 1, 0);
    for (int64_t r = 0; r < a.nrows; ++r) {
        for (int64_t p = a.rowptr[r]; p < a.rowptr[r + 1]; ++p)
            if (a.val[p] != 0.0) {
                a.col[w] = a.col[p];
                a.val[w] = a.val[p];
                ++w;
            }
        rp[r + 1] = w;
    }
    a.col.resize(w);
    a.val.resize(w);
    a.rowptr = rp;
}

void csr_spmv(const Csr& a, const double* x, double* y) {
    for (int64_t r = 0; r < a.nrows; ++r) {
        double s = 0.0;
        for (int64_t p = a.rowptr[r]; p < a.rowptr[r + 1]; ++p) s += a.val[p] * x[a.col[p]];
        y[r] = s;
    }
}
