/*
 * oracle_ops.h -- body of the CPU oracle, instantiated twice by oracle.c
 * (REAL = double with suffix _f64, REAL = float with suffix _f32).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain restatement of RoughPy's dense
 * free-tensor loops, kept in the reference's own loop order so that its
 * rounding follows the reference.  Nothing in the product links it; only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * Layout (roughpy_compute/README.md:110-134, tensor_basis.c:42-58): a dense
 * tensor is one flat array; level k occupies [db[k], db[k+1]) with db[0]=0,
 * db[k+1] = 1 + w*db[k]; the word (i_1..i_k) (0-based letters, first letter
 * most significant) sits at db[k] + sum_j i_j w^(k-j).
 */

/* C += op(A (x) B) with per-operand degree bounds.
 * Restates roughpy_compute/dense/basic/free_tensor_fma.hpp:13-53
 * (out degree D->1, lhs degree hi->lo, i outer, j inner, unit term last).
 * op(v) = v * factor. */
void SFX(orc_ft_fma)(REAL *out, const REAL *lhs, const REAL *rhs, int w,
                     int out_min, int out_max, int lhs_min, int lhs_max,
                     int rhs_min, int rhs_max, REAL factor)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    int Dmax = out_max > lhs_max ? out_max : lhs_max;
    if (rhs_max > Dmax) Dmax = rhs_max;
    orc_degree_begin(w, Dmax, db);
    int omin = out_min > 1 ? out_min : 1;
    for (int od = out_max; od >= omin; --od) {
        int lmax = lhs_max < od - rhs_min ? lhs_max : od - rhs_min;
        int lmin = lhs_min > od - rhs_max ? lhs_min : od - rhs_max;
        REAL *o = out + db[od];
        for (int ld = lmax; ld >= lmin; --ld) {
            int rd = od - ld;
            const REAL *l = lhs + db[ld];
            const REAL *r = rhs + db[rd];
            int64_t ls = db[ld + 1] - db[ld];
            int64_t rs = db[rd + 1] - db[rd];
            for (int64_t i = 0; i < ls; ++i)
                for (int64_t j = 0; j < rs; ++j)
                    o[i * rs + j] += (l[i] * r[j]) * factor;
        }
    }
    if (out_min == 0 && lhs_min == 0 && rhs_min == 0)
        out[0] += (lhs[0] * rhs[0]) * factor;
}

/* A <- op(A (x) B), restating free_tensor_inplace_mul.hpp:13-87: top-down
 * over output degree so lower levels are read before they are overwritten;
 * rhs_min > 0 means an implicitly zero unit.  op(v) = v * factor. */
void SFX(orc_ft_inplace_mul)(REAL *lhs, const REAL *rhs, int w, int lhs_min,
                             int lhs_max, int rhs_min, int rhs_max,
                             REAL factor)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    int Dmax = lhs_max > rhs_max ? lhs_max : rhs_max;
    orc_degree_begin(w, Dmax, db);
    int lo = lhs_min > 1 ? lhs_min : 1;
    for (int od = lhs_max; od >= lo; --od) {
        int lmin = lhs_min > od - rhs_max ? lhs_min : od - rhs_max;
        int lmax = od - (rhs_min > 1 ? rhs_min : 1);
        REAL *o = lhs + db[od];
        int64_t os = db[od + 1] - db[od];
        if (rhs_min == 0) {
            for (int64_t i = 0; i < os; ++i) o[i] = (o[i] * rhs[0]) * factor;
        } else {
            for (int64_t i = 0; i < os; ++i) o[i] = 0;
        }
        for (int ld = lmax; ld >= lmin; --ld) {
            int rd = od - ld;
            const REAL *l = lhs + db[ld];
            const REAL *r = rhs + db[rd];
            int64_t ls = db[ld + 1] - db[ld];
            int64_t rs = db[rd + 1] - db[rd];
            for (int64_t i = 0; i < ls; ++i)
                for (int64_t j = 0; j < rs; ++j)
                    o[i * rs + j] += (l[i] * r[j]) * factor;
        }
    }
    if (lhs_min == 0 && rhs_min == 0)
        lhs[0] = (lhs[0] * rhs[0]) * factor;
    else if (lhs_min == 0)
        lhs[0] = 0;
}

/* exp(X), restating roughpy_compute/dense/intermediate/free_tensor_exp.hpp:
 * 13-37 (Horner; X[0] ignored; DivideBy multiplies by the reciprocal,
 * common/operations.hpp:47-54). */
void SFX(orc_ft_exp)(REAL *out, const REAL *arg, int w, int D)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    orc_degree_begin(w, D, db);
    for (int64_t i = 0; i < db[D + 1]; ++i) out[i] = 0;
    out[0] = 1;
    for (int deg = D; deg > 0; --deg) {
        int ml = D - deg + 1;
        REAL f = (REAL)1 / (REAL)deg;
        SFX(orc_ft_inplace_mul)(out, arg, w, 0, ml, 1, ml, f);
        out[0] += 1;
    }
}

/* A exp(X), restating free_tensor_fmexp.hpp:15-40 (the Chen step). */
void SFX(orc_ft_fmexp)(REAL *out, const REAL *mul, const REAL *expo, int w,
                       int D)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    orc_degree_begin(w, D, db);
    int64_t n = db[D + 1];
    for (int64_t i = 0; i < n; ++i) out[i] = mul[i];
    for (int deg = D; deg > 0; --deg) {
        int ml = D - deg + 1;
        REAL f = (REAL)1 / (REAL)deg;
        SFX(orc_ft_inplace_mul)(out, expo, w, 0, ml, 1, ml, f);
        for (int64_t i = 0; i < n; ++i) out[i] += mul[i];
    }
}

/* log(X) assuming X[0] == 1, restating free_tensor_log.hpp:12-40
 * (out starts at zero as allocated by roughpy/compute/__init__.py:267). */
void SFX(orc_ft_log)(REAL *out, const REAL *arg, int w, int D)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    orc_degree_begin(w, D, db);
    for (int64_t i = 0; i < db[D + 1]; ++i) out[i] = 0;
    for (int deg = D; deg > 0; --deg) {
        int ml = D - deg + 1;
        if (deg % 2 == 0)
            out[0] -= (REAL)1 / (REAL)deg;
        else
            out[0] += (REAL)1 / (REAL)deg;
        SFX(orc_ft_inplace_mul)(out, arg, w, 0, ml, 1, ml, (REAL)1);
    }
}

/* Antipode: reversed word, sign (-1)^deg.  Restates
 * free_tensor_antipode.hpp:45-100 with DefaultSigner. */
void SFX(orc_antipode)(REAL *out, const REAL *arg, int w, int D)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    orc_degree_begin(w, D, db);
    out[0] = arg[0];
    for (int d = 1; d <= D; ++d) {
        int64_t sz = db[d + 1] - db[d];
        for (int64_t i = 0; i < sz; ++i) {
            int64_t a = i, r = 0;
            for (int k = 0; k < d; ++k) { r = r * w + a % w; a /= w; }
            REAL v = arg[db[d] + i];
            out[db[d] + r] = (d % 2) ? -v : v;
        }
    }
}

/* Adjoint of left multiplication by `op`, restating
 * free_tensor_adjoint_left_mul.hpp:10-77:
 *   out_k[i] += sum_{op word u} op[u] * arg[u i]. */
void SFX(orc_adj_lmul)(REAL *out, const REAL *op, const REAL *arg, int w,
                       int out_min, int out_max, int op_min, int op_max,
                       int arg_min, int arg_max)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    int Dmax = out_max > arg_max ? out_max : arg_max;
    if (op_max > Dmax) Dmax = op_max;
    orc_degree_begin(w, Dmax, db);
    int amax = arg_max - op_min < out_max ? arg_max - op_min : out_max;
    int amin = arg_min - op_max > out_min ? arg_min - op_max : out_min;
    for (int ad = amax; ad >= amin; --ad) {
        int omin = ad - op_max > out_min ? ad - op_max : out_min;
        int omax = ad - op_min < out_max ? ad - op_min : out_max;
        const REAL *a = arg + db[ad];
        for (int od = omax; od >= omin; --od) {
            int pd = ad - od;
            const REAL *p = op + db[pd];
            REAL *o = out + db[od];
            int64_t ps = db[pd + 1] - db[pd];
            int64_t os = db[od + 1] - db[od];
            for (int64_t pi = 0; pi < ps; ++pi) {
                REAL pv = p[pi];
                for (int64_t i = 0; i < os; ++i) o[i] += pv * a[i + pi * os];
            }
        }
    }
}

/* Shuffle product out += lhs (sh) rhs at full depth D, restating the v1
 * kernel of roughpy_compute/dense/basic/shuffle_tensor_product.hpp:24-85:
 * out[0] += lhs[0] rhs[0]; then for each output word (degree 1..D, letters
 * unpacked least significant first) the sum over all 2^n bit masks, bit j
 * sending letter j (j = 0 the last letter) to the lhs word and the others to
 * the rhs word (basis.hpp:50-61), accumulated from zero and then added. */
void SFX(orc_st_fma)(REAL *out, const REAL *lhs, const REAL *rhs, int w, int D)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    int letters[ORC_MAX_DEPTH];
    orc_degree_begin(w, D, db);
    out[0] += lhs[0] * rhs[0];
    for (int n = 1; n <= D; ++n) {
        int64_t size = db[n + 1] - db[n];
        for (int64_t i = 0; i < size; ++i) {
            int64_t t = i;
            for (int d = 0; d < n; ++d) {
                letters[d] = (int)(t % w);
                t /= w;
            }
            REAL acc = 0;
            for (uint64_t mask = 0; mask < ((uint64_t)1 << n); ++mask) {
                int64_t li = 0, ri = 0;
                int ld = 0, rd = 0;
                for (int d = n - 1; d >= 0; --d) {
                    if ((mask >> d) & 1) {
                        ++ld;
                        li = li * w + letters[d];
                    } else {
                        ++rd;
                        ri = ri * w + letters[d];
                    }
                }
                acc += lhs[db[ld] + li] * rhs[db[rd] + ri];
            }
            out[db[n] + i] += acc;
        }
    }
}

/* Adjoint of shuffle multiplication by `op`, restating
 * roughpy_compute/dense/basic/shuffle_tensor_adjoint_mul.hpp:13-70: for
 * every word u of arg (degree 0..D) and every bit mask, the masked letters
 * form the op word and the rest the out word:
 *   out[rest] += arg[u] * op[masked]. */
void SFX(orc_st_adj_mul)(REAL *out, const REAL *op, const REAL *arg, int w, int D)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    int letters[ORC_MAX_DEPTH];
    orc_degree_begin(w, D, db);
    for (int n = 0; n <= D; ++n) {
        int64_t size = db[n + 1] - db[n];
        for (int64_t i = 0; i < size; ++i) {
            int64_t t = i;
            for (int d = 0; d < n; ++d) {
                letters[d] = (int)(t % w);
                t /= w;
            }
            REAL scalar = arg[db[n] + i];
            for (uint64_t mask = 0; mask < ((uint64_t)1 << n); ++mask) {
                int64_t pi = 0, oi = 0;
                int pd = 0, od = 0;
                for (int d = n - 1; d >= 0; --d) {
                    if ((mask >> d) & 1) {
                        ++pd;
                        pi = pi * w + letters[d];
                    } else {
                        ++od;
                        oi = oi * w + letters[d];
                    }
                }
                out[db[od] + oi] += scalar * op[db[pd] + pi];
            }
        }
    }
}

/* Signature of a piecewise-linear stream by Chen's identity, restating
 * algebra/src/lite_context.h:545-563: result = 1; for each increment row,
 * result = result * exp(l2t(row)), the exponential applied with the fmexp
 * Horner form (free_tensor_fmexp.hpp:15-40).  l2t of a degree-one Lie
 * element is the identity onto level one (lie_basis.c:782-797).
 * scratch: 3 tensors. */
void SFX(orc_signature)(REAL *out, const REAL *incr, int64_t n, int w, int D,
                        REAL *scratch)
{
    int64_t db[ORC_MAX_DEPTH + 2];
    orc_degree_begin(w, D, db);
    int64_t sz = db[D + 1];
    REAL *x = scratch, *cur = scratch + sz, *nxt = scratch + 2 * sz;
    for (int64_t i = 0; i < sz; ++i) { x[i] = 0; cur[i] = 0; }
    cur[0] = 1;
    for (int64_t s = 0; s < n; ++s) {
        for (int a = 0; a < w; ++a) x[1 + a] = incr[s * w + a];
        SFX(orc_ft_fmexp)(nxt, cur, x, w, D);
        REAL *t = cur; cur = nxt; nxt = t;
    }
    for (int64_t i = 0; i < sz; ++i) out[i] = cur[i];
}

struct SFX(orc_sigjob) {
    REAL *out; const REAL *incr; const int64_t *offsets;
    int64_t first, last; int w, D;
};

static void *SFX(orc_sig_worker)(void *p)
{
    struct SFX(orc_sigjob) *j = (struct SFX(orc_sigjob) *)p;
    int64_t db[ORC_MAX_DEPTH + 2];
    orc_degree_begin(j->w, j->D, db);
    int64_t sz = db[j->D + 1];
    REAL *scratch = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)sz);
    if (!scratch) return NULL;
    for (int64_t s = j->first; s < j->last; ++s) {
        int64_t b = j->offsets[s], e = j->offsets[s + 1];
        SFX(orc_signature)(j->out + s * sz, j->incr + b * j->w, e - b, j->w,
                           j->D, scratch);
    }
    free(scratch);
    return NULL;
}

/* Batched signatures of uneven streams (CSR offsets over increment rows),
 * one stream per thread slice; nthreads <= 1 runs on the calling thread.
 * Mirrors the GIL-free per-stream Stream.signature usage
 * (roughpy/src/streams/stream.cpp:196). */
int SFX(orc_signature_batch)(REAL *out, const REAL *incr,
                             const int64_t *offsets, int64_t n_streams, int w,
                             int D, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    struct SFX(orc_sigjob) jobs[256];
    int64_t per = (n_streams + nthreads - 1) / nthreads;
    int started = 0;
    for (int t = 0; t < nthreads; ++t) {
        int64_t f = t * per, l = f + per;
        if (f >= n_streams) break;
        if (l > n_streams) l = n_streams;
        jobs[t].out = out; jobs[t].incr = incr; jobs[t].offsets = offsets;
        jobs[t].first = f; jobs[t].last = l; jobs[t].w = w; jobs[t].D = D;
        if (nthreads == 1) {
            SFX(orc_sig_worker)(&jobs[t]);
        } else {
            if (pthread_create(&th[t], NULL, SFX(orc_sig_worker), &jobs[t]))
                return -1;
            ++started;
        }
    }
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* Compressed sparse matrix applied to a dense vector, restating
 * roughpy_compute/dense/basic/apply_sparse_linear_map.hpp:16-55:
 * csr (compressed_rows != 0): out[i] += sum_j c_ij arg[j];
 * csc: out[i] += c_ij arg[j] scattered column by column.  op = * scale. */
void SFX(orc_sparse_apply)(REAL *out, const REAL *arg, const int64_t *offsets,
                           const int64_t *indices, const REAL *data,
                           int64_t n_offsets, int compressed_rows, REAL scale)
{
    if (compressed_rows) {
        for (int64_t i = 0; i < n_offsets; ++i)
            for (int64_t p = offsets[i]; p < offsets[i + 1]; ++p)
                out[i] += (data[p] * arg[indices[p]]) * scale;
    } else {
        for (int64_t j = 0; j < n_offsets; ++j)
            for (int64_t p = offsets[j]; p < offsets[j + 1]; ++p)
                out[indices[p]] += (data[p] * arg[j]) * scale;
    }
}
