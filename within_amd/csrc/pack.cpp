// pack.cpp -- independent problems of one shape packed into one block-diagonal design.
//
// Member m's observations keep their own internal order and occupy the merged rows
// [row_off, row_off + n).  Term t of the merged design concatenates the members' term-t levels,
// each member's range padded to PACK_ALIGN levels, so every workgroup of a DOF-space kernel
// (at most KERNEL_BLOCK DOFs) lies inside one member.  Codes become merged level indices, the
// preconditioner data are concatenated with their DOF indices remapped, and the fused term's
// work items are padded per member with empty items the same way.  Padding levels have no
// observations, are covered by no subdomain and stay zero through the solve.
#include <algorithm>
#include <atomic>
#include <list>
#include <mutex>

#include "schwarz.hpp"
#include "solver.hpp"

namespace wx {

uint64_t next_solver_uid() {
    static std::atomic<uint64_t> counter{1};
    return counter.fetch_add(1);
}

namespace {

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

int precond_kind(const Solver &s) {
    if (!s.precond) return 0;
    return s.precond->kind == Precond::DIAGONAL ? 1 : 2;
}

// member DOF d (its own term-major layout) -> merged DOF
struct DofMap {
    uint32_t T = 0;
    uint64_t own_off[MAX_TERMS] = {}, L[MAX_TERMS] = {}, merged_off[MAX_TERMS] = {};
    uint64_t operator()(uint64_t d) const {
        uint32_t t = 0;
        while (t + 1 < T && d >= own_off[t + 1]) ++t;
        return merged_off[t] + (d - own_off[t]);
    }
};

std::shared_ptr<Precond> merge_schwarz(const std::vector<const Solver *> &ms, const std::vector<DofMap> &maps,
                                       uint64_t n_dofs) {
    auto p = std::make_shared<Precond>();
    p->kind = Precond::ADDITIVE;
    p->n = n_dofs;
    auto S = std::make_shared<Precond::Schwarz>();
    S->n_dofs = n_dofs;
    S->slot_start.push_back(0);
    S->c_ptr.push_back(0);
    for (size_t m = 0; m < ms.size(); ++m) {
        const Precond::Schwarz &A = *ms[m]->precond->schwarz;
        const uint64_t slot_base = S->glob.size(), c_base = S->c_idx.size(), dense_base = S->dense.size(),
                       lptr_base = S->l_ptr.size(), lidx_base = S->l_idx.size(), piv_base = S->piv.size(),
                       cover_base = S->cover_perm.size();
        for (uint64_t c = 0; c < A.n_comp(); ++c) {
            S->slot_start.push_back(slot_base + A.slot_start[c + 1]);
            S->ne.push_back(A.ne[c]);
            S->nk.push_back(A.nk[c]);
            S->fac_kind.push_back(A.fac_kind[c]);
            S->gauge.push_back(A.gauge[c]);
            S->dense_off.push_back(dense_base + A.dense_off[c]);
            S->l_row_off.push_back(lptr_base + A.l_row_off[c]);
            S->piv_off.push_back(piv_base + A.piv_off[c]);
            S->cover_off.push_back(cover_base + A.cover_off[c]);
        }
        for (uint64_t s = 0; s < A.n_slots(); ++s) {
            S->glob.push_back((uint32_t)maps[m](A.glob[s]));
            S->c_ptr.push_back(c_base + A.c_ptr[s + 1]);
        }
        S->pw.insert(S->pw.end(), A.pw.begin(), A.pw.end());
        S->fold.insert(S->fold.end(), A.fold.begin(), A.fold.end());
        S->inv_de.insert(S->inv_de.end(), A.inv_de.begin(), A.inv_de.end());
        S->c_idx.insert(S->c_idx.end(), A.c_idx.begin(), A.c_idx.end());
        S->c_val.insert(S->c_val.end(), A.c_val.begin(), A.c_val.end());
        A.host_dense();
        S->dense.insert(S->dense.end(), A.dense.begin(), A.dense.end());
        for (uint64_t v : A.l_ptr) S->l_ptr.push_back(lidx_base + v);
        S->l_idx.insert(S->l_idx.end(), A.l_idx.begin(), A.l_idx.end());
        S->l_val.insert(S->l_val.end(), A.l_val.begin(), A.l_val.end());
        S->piv.insert(S->piv.end(), A.piv.begin(), A.piv.end());
        S->cover_perm.insert(S->cover_perm.end(), A.cover_perm.begin(), A.cover_perm.end());
    }
    p->schwarz = S;
    return p;
}

std::shared_ptr<Pack> build_pack(const std::vector<const Solver *> &ms, const std::vector<uint32_t> &kv, uint32_t tile) {
    const uint32_t k = *std::max_element(kv.begin(), kv.end());
    auto pk = std::make_shared<Pack>();
    pk->k = k;
    pk->kshift = pow2_shift(k);
    const uint32_t g = KERNEL_BLOCK >> pk->kshift; // DOFs and work items per workgroup
    const size_t M = ms.size();
    const Design &d0 = ms[0]->design;
    const uint32_t T = (uint32_t)d0.terms.size();
    const bool weighted = ms[0]->weighted;
    // level layout
    std::vector<uint64_t> Lt(T, 0);
    pk->members.resize(M);
    uint64_t rows = 0;
    uint32_t tiles = 0;
    for (size_t m = 0; m < M; ++m) {
        const Design &d = ms[m]->design;
        PackMember &mb = pk->members[m];
        mb.s = ms[m];
        mb.kv = kv[m];
        mb.n = d.n_obs;
        mb.row_off = rows;
        rows += d.n_obs;
        mb.n_dofs = d.n_dofs;
        mb.n_tiles = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((d.n_obs + 63) / 64, MAX_PART_BLOCKS));
        mb.tile_off = tiles;
        tiles += mb.n_tiles;
        for (uint32_t t = 0; t < T; ++t) {
            mb.L[t] = d.terms[t].n_levels;
            mb.own_off[t] = d.terms[t].offset;
            mb.lvl_off[t] = Lt[t];
            Lt[t] += round_up(mb.L[t], PACK_ALIGN);
        }
    }
    if (rows > UINT32_MAX) fail(WX_ERR_INVALID_INPUT, "packed problems exceed 2^32 observations");
    pk->n_tiles = tiles;
    // merged design
    auto sv = std::make_unique<Solver>();
    Design &md = sv->design;
    md.n_obs = rows;
    md.terms.resize(T);
    uint64_t off = 0;
    std::vector<uint64_t> term_off(T);
    for (uint32_t t = 0; t < T; ++t) {
        md.terms[t].n_levels = Lt[t];
        md.terms[t].offset = off;
        md.terms[t].sorted = d0.terms[t].sorted;
        md.terms[t].loads = {-1};
        term_off[t] = off;
        off += Lt[t];
    }
    md.n_dofs = off;
    if (off > UINT32_MAX) fail(WX_ERR_INVALID_INPUT, "packed problems exceed 2^32 coefficients");
    md.codes.resize((size_t)T * rows);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t m = 0; m < (int64_t)M; ++m) {
        const Design &d = ms[m]->design;
        const PackMember &mb = pk->members[m];
        for (uint32_t t = 0; t < T; ++t) {
            const uint32_t *src = d.level_column(t);
            uint32_t *dst = md.codes.data() + (size_t)t * rows + mb.row_off;
            const uint32_t lo = (uint32_t)mb.lvl_off[t];
            for (uint64_t i = 0; i < d.n_obs; ++i) dst[i] = lo + src[i];
        }
    }
    std::vector<double> w;
    if (weighted) {
        w.reserve(rows);
        for (const Solver *s : ms) w.insert(w.end(), s->w_internal.begin(), s->w_internal.end());
        sv->min_weight = *std::min_element(w.begin(), w.end());
    }
    sv->weighted = weighted;
    // preconditioner
    std::vector<DofMap> maps(M);
    for (size_t m = 0; m < M; ++m) {
        maps[m].T = T;
        for (uint32_t t = 0; t < T; ++t) {
            maps[m].own_off[t] = pk->members[m].own_off[t];
            maps[m].L[t] = pk->members[m].L[t];
            maps[m].merged_off[t] = term_off[t] + pk->members[m].lvl_off[t];
        }
    }
    const int kind = precond_kind(*ms[0]);
    if (kind == 1) {
        auto p = std::make_shared<Precond>();
        p->kind = Precond::DIAGONAL;
        p->n = md.n_dofs;
        p->inv_diag.assign(md.n_dofs, 0.0);
        for (size_t m = 0; m < M; ++m)
            for (uint64_t q = 0; q < ms[m]->design.n_dofs; ++q) p->inv_diag[maps[m](q)] = ms[m]->precond->inv_diag[q];
        sv->precond = p;
    } else if (kind == 2) {
        sv->precond = merge_schwarz(ms, maps, md.n_dofs);
    }
    sv->uid = next_solver_uid();
    upload_design(*sv, w, tile);
    DevDesign &dv = sv->dev;
    if (dv.fuse_term < 0) fail(WX_ERR_INVALID_INPUT, "packed design has no sorted term to fuse");
    // fused term: the gather walks one level (or level chunk) per item, each member's items
    // padded to PACK_ALIGN (every workgroup inside one member)
    DevTerm &ft = dv.terms[dv.fuse_term];
    const uint32_t fterm = (uint32_t)dv.fuse_term;
    std::vector<SegItem> items(ft.n_items), padded;
    WX_HIP(hipMemcpy(items.data(), ft.items.p, items.size() * sizeof(SegItem), hipMemcpyDeviceToHost));
    padded.reserve(items.size() + M * PACK_ALIGN);
    std::vector<uint32_t> item_rng, item_ptr{0};
    size_t q = 0;
    for (size_t m = 0; m < M; ++m) {
        const uint64_t lo = pk->members[m].lvl_off[fterm], hi = lo + round_up(pk->members[m].L[fterm], PACK_ALIGN);
        const uint64_t b = padded.size();
        while (q < items.size() && items[q].level >= lo && items[q].level < hi) padded.push_back(items[q++]);
        while (padded.size() % PACK_ALIGN) padded.push_back({0u, 0u, 0u, SEG_SKIP});
        item_rng.push_back((uint32_t)(b / g));
        item_rng.push_back((uint32_t)(padded.size() / g));
        item_ptr.push_back((uint32_t)(m + 1));
    }
    if (q != items.size()) fail(WX_ERR_INVALID_INPUT, "packed design: work items out of member order");
    ft.gitems.upload(padded.data(), padded.size());
    ft.n_gitems = (uint32_t)padded.size();
    pk->n_item_blocks = (uint32_t)(padded.size() / g);
    std::vector<uint32_t> item_col(pk->n_item_blocks);
    for (size_t m = 0; m < M; ++m)
        for (uint32_t b = item_rng[2 * m]; b < item_rng[2 * m + 1]; ++b) item_col[b] = (uint32_t)(m * k);
    // DOF blocks, their members and ranges; the own index of every merged DOF
    const uint64_t Nm = md.n_dofs;
    pk->n_dof_blocks = (uint32_t)(Nm / g);
    std::vector<uint32_t> dof_member(Nm, 0);
    std::vector<int32_t> dof_own(Nm, -1);
    std::vector<uint32_t> dof_rng, dof_ptr{0};
    std::vector<uint64_t> own_ld(M);
    std::vector<uint32_t> own_kv(M);
    for (size_t m = 0; m < M; ++m) {
        const PackMember &mb = pk->members[m];
        own_ld[m] = mb.n_dofs;
        own_kv[m] = mb.kv;
        for (uint32_t t = 0; t < T; ++t) {
            const uint64_t b = term_off[t] + mb.lvl_off[t], len = round_up(mb.L[t], PACK_ALIGN);
            for (uint64_t l = 0; l < len; ++l) {
                dof_member[b + l] = (uint32_t)m;
                if (l < mb.L[t]) dof_own[b + l] = (int32_t)(mb.own_off[t] + l);
            }
            dof_rng.push_back((uint32_t)(b / g));
            dof_rng.push_back((uint32_t)((b + len) / g));
        }
        dof_ptr.push_back((uint32_t)((m + 1) * T));
    }
    std::vector<uint32_t> dof_col(pk->n_dof_blocks);
    for (uint32_t b = 0; b < pk->n_dof_blocks; ++b) dof_col[b] = dof_member[(uint64_t)b * g] * k;
    std::vector<uint32_t> tile_rng, tile_ptr{0};
    for (size_t m = 0; m < M; ++m) {
        tile_rng.push_back(pk->members[m].tile_off);
        tile_rng.push_back(pk->members[m].tile_off + pk->members[m].n_tiles);
        tile_ptr.push_back((uint32_t)(m + 1));
    }
    pk->dof_col.upload(dof_col.data(), dof_col.size());
    pk->item_col.upload(item_col.data(), item_col.size());
    pk->dof_rng_ptr.upload(dof_ptr.data(), dof_ptr.size());
    pk->dof_rng.upload(dof_rng.data(), dof_rng.size());
    pk->item_rng_ptr.upload(item_ptr.data(), item_ptr.size());
    pk->item_rng.upload(item_rng.data(), item_rng.size());
    pk->tile_rng_ptr.upload(tile_ptr.data(), tile_ptr.size());
    pk->tile_rng.upload(tile_rng.data(), tile_rng.size());
    pk->dof_member.upload(dof_member.data(), dof_member.size());
    pk->dof_own.upload(dof_own.data(), dof_own.size());
    pk->own_ld.upload(own_ld.data(), own_ld.size());
    pk->own_kv.upload(own_kv.data(), own_kv.size());
    pk->part_blocks = std::max<uint64_t>({pk->n_dof_blocks, pk->n_item_blocks, pk->n_tiles});
    pk->merged = std::move(sv);
    return pk;
}

std::mutex g_pack_mu;
std::list<std::pair<std::vector<uint64_t>, std::shared_ptr<Pack>>> g_packs; // most recent first
constexpr size_t PACK_CACHE = 16;

} // namespace

bool pack_signature(const Solver &s, std::string &sig) {
    const Design &d = s.design;
    if (d.has_slopes() || s.reparam || !s.dev.plain || s.dev.fuse_term < 0 || d.terms.size() > MAX_TERMS)
        return false;
    if (s.weighted && s.w_internal.size() != d.n_obs) return false;
    sig = "T" + std::to_string(d.terms.size()) + (s.weighted ? "w" : "u") + std::to_string(precond_kind(s)) + "s";
    for (const Term &t : d.terms) sig += t.sorted ? '1' : '0';
    return true;
}

std::shared_ptr<Pack> get_pack(const std::vector<const Solver *> &members, const std::vector<uint32_t> &kv,
                               uint32_t tile) {
    std::vector<uint64_t> key{tile};
    for (size_t m = 0; m < members.size(); ++m) {
        key.push_back(members[m]->uid);
        key.push_back(kv[m]);
    }
    {
        std::lock_guard<std::mutex> lk(g_pack_mu);
        for (auto it = g_packs.begin(); it != g_packs.end(); ++it)
            if (it->first == key) {
                g_packs.splice(g_packs.begin(), g_packs, it);
                return it->second;
            }
    }
    auto pk = build_pack(members, kv, tile);
    std::lock_guard<std::mutex> lk(g_pack_mu);
    g_packs.emplace_front(key, pk);
    while (g_packs.size() > PACK_CACHE) g_packs.pop_back();
    return pk;
}

} // namespace wx
