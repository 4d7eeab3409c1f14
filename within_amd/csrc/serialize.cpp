// serialize.cpp -- versioned binary persistence of a preconditioner.
//
// Replaces the postcard wire format of the reference's Preconditioner (crates/within-py/src/
// config.rs:367-388, wire v13).  Byte compatibility with postcard is not attempted: the
// reference's factor layout belongs to an external crate.  The format is little-endian:
//   magic "WXPC" | u32 version | u32 kind | u64 n | payload
// Diagonal payload: f64 vector inv_diag.  Additive payload: every array of
// Precond::Schwarz as (u64 length, raw elements), in declaration order.  Decoding validates every length and
// index against the others, so corrupt or truncated bytes raise an error, never a fault.
#include <cmath>
#include <cstring>

#include "schwarz.hpp"

namespace wx {

namespace {
constexpr uint32_t VERSION = 2;

struct Writer {
    std::vector<uint8_t> b;
    template <typename T> void put(T v) {
        const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    template <typename V> void vec(const V &v) {
        put<uint64_t>(v.size());
        const uint8_t *p = reinterpret_cast<const uint8_t *>(v.data());
        b.insert(b.end(), p, p + v.size() * sizeof(typename V::value_type));
    }
};

struct Reader {
    const uint8_t *p;
    uint64_t len, pos = 0;
    template <typename T> T get() {
        if (len - pos < sizeof(T)) fail(WX_ERR_INVALID_INPUT, "preconditioner bytes are truncated");
        T v;
        std::memcpy(&v, p + pos, sizeof(T));
        pos += sizeof(T);
        return v;
    }
    template <typename T> std::vector<T> vec() {
        const uint64_t n = get<uint64_t>();
        if (n > (len - pos) / sizeof(T)) fail(WX_ERR_INVALID_INPUT, "preconditioner bytes are truncated");
        std::vector<T> v(n);
        if (n) std::memcpy(v.data(), p + pos, n * sizeof(T));
        pos += n * sizeof(T);
        return v;
    }
};

[[noreturn]] void bad(const char *what) { fail(WX_ERR_INVALID_INPUT, std::string("invalid preconditioner bytes: ") + what); }
} // namespace

std::vector<uint8_t> serialize_precond(const Precond &p) {
    Writer w;
    w.b.insert(w.b.end(), {'W', 'X', 'P', 'C'});
    w.put<uint32_t>(VERSION);
    w.put<uint32_t>((uint32_t)p.kind);
    w.put<uint64_t>(p.n);
    if (p.kind == Precond::DIAGONAL) {
        w.vec(p.inv_diag);
    } else {
        const auto &S = *p.schwarz;
        w.vec(S.slot_start);
        w.vec(S.ne);
        w.vec(S.nk);
        w.vec(S.fac_kind);
        w.vec(S.gauge);
        w.vec(S.glob);
        w.vec(S.pw);
        w.vec(S.fold);
        w.vec(S.inv_de);
        w.vec(S.c_ptr);
        w.vec(S.c_idx);
        w.vec(S.c_val);
        w.vec(S.dense_off);
        S.host_dense();
        w.vec(S.dense);
        w.vec(S.l_row_off);
        w.vec(S.l_ptr);
        w.vec(S.l_idx);
        w.vec(S.l_val);
        w.vec(S.piv_off);
        w.vec(S.piv);
        w.vec(S.cover_off);
        w.vec(S.cover_perm);
    }
    return w.b;
}

std::shared_ptr<Precond> deserialize_precond(const uint8_t *buf, uint64_t len) {
    Reader r{buf, len};
    if (len < 4 || std::memcmp(buf, "WXPC", 4) != 0) bad("missing magic");
    r.pos = 4;
    const uint32_t ver = r.get<uint32_t>();
    if (ver != VERSION) bad("unsupported format version");
    const uint32_t kind = r.get<uint32_t>();
    const uint64_t n = r.get<uint64_t>();
    auto p = std::make_shared<Precond>();
    p->n = n;
    if (kind == Precond::DIAGONAL) {
        p->kind = Precond::DIAGONAL;
        p->inv_diag = r.vec<double>();
        if (p->inv_diag.size() != n) bad("diagonal length disagrees with n");
        for (double v : p->inv_diag)
            if (!std::isfinite(v)) bad("non-finite diagonal entry");
    } else if (kind == Precond::ADDITIVE) {
        p->kind = Precond::ADDITIVE;
        auto S = std::make_shared<Precond::Schwarz>();
        S->n_dofs = n;
        S->slot_start = r.vec<uint64_t>();
        S->ne = r.vec<uint32_t>();
        S->nk = r.vec<uint32_t>();
        S->fac_kind = r.vec<uint8_t>();
        S->gauge = r.vec<uint8_t>();
        S->glob = r.vec<uint32_t>();
        S->pw = r.vec<double>();
        S->fold = r.vec<double>();
        S->inv_de = r.vec<double>();
        S->c_ptr = r.vec<uint64_t>();
        S->c_idx = r.vec<uint32_t>();
        S->c_val = r.vec<double>();
        S->dense_off = r.vec<uint64_t>();
        {
            const std::vector<double> dn = r.vec<double>();
            S->dense.assign(dn.begin(), dn.end());
        }
        S->l_row_off = r.vec<uint64_t>();
        S->l_ptr = r.vec<uint64_t>();
        S->l_idx = r.vec<uint32_t>();
        S->l_val = r.vec<double>();
        S->piv_off = r.vec<uint64_t>();
        S->piv = r.vec<double>();
        S->cover_off = r.vec<uint64_t>();
        S->cover_perm = r.vec<uint32_t>();
        if (r.pos != len) bad("trailing bytes");
        // cross-field validation
        const uint64_t nc = S->ne.size();
        if (nc == 0) bad("no components");
        if (S->nk.size() != nc || S->fac_kind.size() != nc || S->gauge.size() != nc || S->slot_start.size() != nc + 1 ||
            S->dense_off.size() != nc || S->l_row_off.size() != nc || S->piv_off.size() != nc ||
            S->cover_off.size() != nc)
            bad("component table lengths disagree");
        for (uint8_t g : S->gauge)
            if (g > 1) bad("unknown gauge");
        if (S->slot_start[0] != 0) bad("slot table does not start at 0");
        for (uint64_t c = 0; c < nc; ++c)
            if (S->slot_start[c + 1] < S->slot_start[c] ||
                S->slot_start[c + 1] - S->slot_start[c] != (uint64_t)S->ne[c] + S->nk[c])
                bad("slot ranges disagree with block sizes");
        const uint64_t ns = S->slot_start[nc];
        if (S->glob.size() != ns || S->pw.size() != ns || S->fold.size() != ns || S->inv_de.size() != ns ||
            S->c_ptr.size() != ns + 1)
            bad("slot array lengths disagree");
        for (double v : S->fold)
            if (!std::isfinite(v) || v == 0.0) bad("invalid coordinate factor");
        for (uint32_t g : S->glob)
            if (g >= n) bad("global index out of range");
        if (S->c_idx.size() != S->c_val.size() || S->c_ptr[ns] != S->c_idx.size() || S->c_ptr[0] != 0)
            bad("C structure invalid");
        for (uint64_t s = 0; s < ns; ++s)
            if (S->c_ptr[s + 1] < S->c_ptr[s]) bad("C row pointers decrease");
        for (uint64_t c = 0; c < nc; ++c) {
            const uint64_t base = S->slot_start[c];
            for (uint64_t s = base; s < base + S->ne[c] + S->nk[c]; ++s)
                for (uint64_t q = S->c_ptr[s]; q < S->c_ptr[s + 1]; ++q)
                    if (s >= base + S->ne[c] || S->c_idx[q] >= S->nk[c]) bad("C entry out of its block");
            if (S->fac_kind[c] == 0) {
                const uint64_t nk = S->nk[c];
                if (S->dense_off[c] > S->dense.size() || S->dense.size() - S->dense_off[c] < nk * nk)
                    bad("dense factor out of range");
            } else if (S->fac_kind[c] == 1 || S->fac_kind[c] == 2) {
                // factor rows: the kept block, or the two sheets of its Gremban cover
                const uint64_t nf = (uint64_t)S->nk[c] * (S->fac_kind[c] == 2 ? 2 : 1), ro = S->l_row_off[c];
                if (ro > S->l_ptr.size() || S->l_ptr.size() - ro < nf + 1) bad("sparse factor rows out of range");
                if (S->piv_off[c] > S->piv.size() || S->piv.size() - S->piv_off[c] < nf) bad("pivots out of range");
                for (uint64_t i = 0; i < nf; ++i) {
                    const uint64_t b = S->l_ptr[ro + i], e = S->l_ptr[ro + i + 1];
                    if (e < b || e > S->l_idx.size()) bad("sparse factor row pointers invalid");
                    for (uint64_t q = b; q < e; ++q)
                        if (S->l_idx[q] >= i) bad("sparse factor entry not strictly lower");
                }
                if (S->fac_kind[c] == 2) {
                    if (S->gauge[c] != 1) bad("a cover component carries no gauge");
                    const uint64_t co = S->cover_off[c];
                    if (co > S->cover_perm.size() || S->cover_perm.size() - co < nf) bad("cover order out of range");
                    std::vector<uint8_t> seen(nf, 0);
                    for (uint64_t t = 0; t < nf; ++t) {
                        const uint32_t v = S->cover_perm[co + t];
                        if (v >= nf || seen[v]) bad("cover order is not a permutation");
                        seen[v] = 1;
                    }
                }
            } else bad("unknown factor kind");
        }
        if (S->l_idx.size() != S->l_val.size()) bad("sparse factor lengths disagree");
        for (double v : S->pw)
            if (!std::isfinite(v)) bad("non-finite partition weight");
        p->schwarz = S;
    } else {
        bad("unknown preconditioner kind");
    }
    return p;
}

} // namespace wx
