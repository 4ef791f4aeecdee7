// C-ABI of markovkit_amd: handles, device residency, the multigrid V-cycle and flexible GMRES
// orchestrated on one HIP stream. Reference constructs replaced by each entry point are cited
// in include/markovkit_amd.h.
#include "runtime.hpp"

#include <map>
#include <mutex>
#include <string>

using namespace mk::rt;

namespace {

// red-black smoother: one row-mapped operator per colour (that colour's rows with all their
// entries, columns unchanged) and the colour of every row
template <typename T>
void split_colours(const mk::generator_csr &cat, const std::vector<std::uint8_t> &colour,
                   level_dev<T> &lv) {
    for (int c = 0; c < 2; ++c) {
        mk::generator_csr sub;
        sub.ptr.assign(1, 0);
        std::vector<u32> rows;
        for (i64 i = 0; i < cat.n; ++i) {
            if (colour[i] != c) continue;
            rows.push_back(static_cast<u32>(i));
            sub.ptr.push_back(sub.ptr.back() + (cat.ptr[i + 1] - cat.ptr[i]));
        }
        sub.n = static_cast<i64>(rows.size());
        sub.col.resize(sub.ptr.back());
        sub.val.resize(sub.ptr.back());
        sub.exit.resize(sub.n);
        const i64 chunk = 1 << 15;
        mk::parallel_tasks((sub.n + chunk - 1) / chunk, [&](i64 t) {
            for (i64 q = t * chunk; q < std::min(sub.n, (t + 1) * chunk); ++q) {
                const i64 i = rows[q];
                std::copy(cat.col.begin() + cat.ptr[i], cat.col.begin() + cat.ptr[i + 1], sub.col.begin() + sub.ptr[q]);
                std::copy(cat.val.begin() + cat.ptr[i], cat.val.begin() + cat.ptr[i + 1], sub.val.begin() + sub.ptr[q]);
                sub.exit[q] = cat.exit[i];
            }
        });
        lv.Ac[c] = upload_csr<T>(sub);
        lv.rowmap[c] = upload(rows);
        lv.Ac[c].view.rowmap = dp<u32>(lv.rowmap[c]);
    }
    lv.colour = upload(colour);
    lv.rb = true;
}

// launches of a level schedule: groups of at least kIluWide rows get a grid each, runs of
// smaller groups share one workgroup
constexpr i64 kIluWide = 2048;

template <typename T>
void ilu_schedule(const std::vector<i64> &depth, i64 groups, dbuf &rows_d, dbuf &ptr_d,
                  std::vector<typename level_dev<T>::ilu_launch> &launches) {
    std::vector<u32> ptr(groups + 1, 0), rows(depth.size());
    for (i64 d : depth) ++ptr[d + 1];
    for (i64 g = 0; g < groups; ++g) ptr[g + 1] += ptr[g];
    std::vector<u32> fill(ptr.begin(), ptr.end() - 1);
    for (std::size_t i = 0; i < depth.size(); ++i) rows[fill[depth[i]]++] = static_cast<u32>(i);
    rows_d = upload(rows);
    ptr_d = upload(ptr);
    launches.clear();
    for (i64 g = 0; g < groups;) {
        const i64 size = ptr[g + 1] - ptr[g];
        if (size >= kIluWide) {
            launches.push_back({static_cast<int>(g), static_cast<int>(g + 1), size});
            ++g;
            continue;
        }
        i64 e = g;
        while (e < groups && ptr[e + 1] - ptr[e] < kIluWide) ++e;
        launches.push_back({static_cast<int>(g), static_cast<int>(e), 0});
        g = e;
    }
}

// ILU(0) of the level operator in rank order (rdme/cpp/rdme_full_multigrid.cpp:208-236, the
// structured case): the graph is bipartite, so only the pivots change,
//   pivot_i = shift + exit_i - sum_{j < i} a_ij a_ji / pivot_j  (a_ij a_ji = rate(j->i) rate(i->j));
// and the dependency depths of the forward (lower neighbours) and backward (upper neighbours)
// solves. Columns of every row are ascending.
template <typename T>
void build_ilu(const mk::generator_csr &cat, double shift, level_dev<T> &lv) {
    const i64 n = cat.n;
    std::vector<double> pivot(n);
    std::vector<T> pinv(n);
    std::vector<i64> fdepth(n, 0), bdepth(n, 0);
    i64 fgroups = n > 0 ? 1 : 0, bgroups = n > 0 ? 1 : 0;
    for (i64 i = 0; i < n; ++i) {
        double d = shift + cat.exit[i];
        for (u64 p = cat.ptr[i]; p < cat.ptr[i + 1] && cat.col[p] < i; ++p) {
            const u32 j = cat.col[p];
            const auto b = cat.col.begin() + cat.ptr[j], e = cat.col.begin() + cat.ptr[j + 1];
            const auto it = std::lower_bound(b, e, static_cast<u32>(i));
            if (it == e || *it != static_cast<u32>(i))
                throw std::logic_error("ILU(0) smoother: a diffusion jump has no reverse jump");
            d -= cat.val[p] * cat.val[cat.ptr[j] + (it - b)] / pivot[j];
            fdepth[i] = std::max(fdepth[i], fdepth[j] + 1);
        }
        if (!(d > 0.0) || !std::isfinite(d)) throw std::runtime_error("population_multigrid: ILU(0) pivot");
        pivot[i] = d;
        pinv[i] = static_cast<T>(1.0 / d);
        fgroups = std::max(fgroups, fdepth[i] + 1);
    }
    for (i64 i = n; i-- > 0;) {
        for (u64 p = cat.ptr[i + 1]; p-- > cat.ptr[i] && cat.col[p] > i;)
            bdepth[i] = std::max(bdepth[i], bdepth[cat.col[p]] + 1);
        bgroups = std::max(bgroups, bdepth[i] + 1);
    }
    if (fgroups >= (i64{1} << 31) || bgroups >= (i64{1} << 31)) throw std::overflow_error("ILU(0) schedule too deep");
    lv.pinv = upload(pinv);
    ilu_schedule<T>(fdepth, fgroups, lv.fw_rows, lv.fw_ptr, lv.fw_launch);
    ilu_schedule<T>(bdepth, bgroups, lv.bw_rows, lv.bw_ptr, lv.bw_launch);
    lv.ilu = true;
}

// the tail's LDS image in HBM: every constant array copied to its arena offset (again after the
// terminal blocks change with the shift)
void refresh_tail_image(mk_sector_mg_s &mg) {
    if (mg.tail_lds_bytes == 0) return;
    for (const auto &cp : mg.tail_copies)
        HIPCHK(hipMemcpyAsync(static_cast<char *>(mg.tail_image.p) + cp.offset, cp.src, cp.bytes,
                              hipMemcpyDeviceToDevice, stream()));
    drain();
}

// LDS layout of the fused tail: every array of levels tail_start .. end at a 16-byte aligned
// offset of one arena, the constant ones first (their image is assembled in HBM) and the work
// vectors after them; disabled when an operator is not coded or the arena exceeds the LDS limit
template <typename T>
void build_tail_lds(mk_sector_mg_s &mg, const std::vector<level_dev<T>> &L,
                    const std::vector<gpu::level_desc<T>> &desc) {
    mg.tail_lds_bytes = 0;
    mg.tail_image_bytes = 0;
    mg.tail_copies.clear();
    if (mg.tail_start < 0) return;
    const i64 depth = static_cast<i64>(L.size());
    const int levels = static_cast<int>(depth - mg.tail_start);
    std::vector<gpu::level_desc<T>> ld;
    std::vector<gpu::lds_copy> plan;
    // the descriptors first
    const u64 dbytes = (static_cast<u64>(levels) * sizeof(gpu::level_desc<T>) + 15) / 16 * 16;
    u64 off = dbytes;
    // offset + 1 as a pointer (0 stays null); src null: a work array, nothing to copy
    auto place = [&](const void *src, u64 bytes, bool needed) -> std::uintptr_t {
        if (!needed || bytes == 0) return 0;
        const u64 o = off;
        off = (off + bytes + 15) / 16 * 16;
        if (src) plan.push_back({src, static_cast<u32>(o), static_cast<u32>(bytes)});
        return static_cast<std::uintptr_t>(o + 1);
    };
    auto view = [&](const dev_csr<T> &A, gpu::csr_view<T> &v) {
        // the image holds dictionary-coded exit rates only (per-row exit arrays stay in HBM)
        if (!A.coded || A.n == 0 || A.view.exit) return A.n == 0;
        const u64 n = static_cast<u64>(A.n);
        v.soff = reinterpret_cast<const u32 *>(place(A.view.soff, (A.view.n_slices + 1) * 4, true));
        v.ent = reinterpret_cast<const u32 *>(place(A.view.ent, A.padded * 4, A.padded > 0));
        v.rlen = reinterpret_cast<const std::uint8_t *>(place(A.view.rlen, n, true));
        v.rcode = reinterpret_cast<const std::uint8_t *>(place(A.view.rcode, n, true));
        v.vdict = reinterpret_cast<const T *>(place(A.view.vdict, A.view.n_vdict * sizeof(T), true));
        v.edict = reinterpret_cast<const T *>(place(A.view.edict, A.view.n_edict * sizeof(T), true));
        return true;
    };
    bool ok = levels <= 12;
    for (i64 l = mg.tail_start; l < depth && ok; ++l) {
        const auto &lv = L[l];
        gpu::level_desc<T> d = desc[l];
        const u64 n = static_cast<u64>(lv.n), ts = sizeof(T);
        ok = view(lv.A, d.A) && (lv.n_next == 0 || view(lv.R, d.R));
        d.coarse_of = reinterpret_cast<const u32 *>(place(desc[l].coarse_of, n * 4, lv.n_next > 0));
        d.weight = reinterpret_cast<const T *>(place(desc[l].weight, n * ts, lv.n_next > 0));
        u64 tinv = 0;
        for (u32 m : lv.h_tsize) tinv += static_cast<u64>(m) * m;
        const bool term = lv.n_term > 0;
        d.tstart = reinterpret_cast<const u32 *>(place(desc[l].tstart, lv.n_term * 4, term));
        d.tsize = reinterpret_cast<const u32 *>(place(desc[l].tsize, lv.n_term * 4, term));
        d.toff = reinterpret_cast<const u64 *>(place(desc[l].toff, lv.n_term * 8, term));
        d.tinv = reinterpret_cast<const T *>(place(desc[l].tinv, tinv * ts, term));
        d.colour = reinterpret_cast<const std::uint8_t *>(place(desc[l].colour, n, desc[l].colour != nullptr));
        for (int c = 0; c < 2; ++c)
            d.crows[c] = reinterpret_cast<const u32 *>(place(desc[l].crows[c], static_cast<u64>(desc[l].ncrows[c]) * 4,
                                                             desc[l].crows[c] != nullptr));
        d.perm = reinterpret_cast<const u32 *>(place(desc[l].perm, n * 4, desc[l].perm != nullptr));
        ld.push_back(d);
    }
    const u64 image_bytes = off;
    for (std::size_t q = 0; q < ld.size() && ok; ++q) {
        const u64 n = static_cast<u64>(ld[q].n), ts = sizeof(T);
        ld[q].acc = reinterpret_cast<T *>(place(nullptr, n * ts, ld[q].acc != nullptr));
        ld[q].rhs = reinterpret_cast<T *>(place(nullptr, n * ts, true));
        ld[q].x = reinterpret_cast<T *>(place(nullptr, n * ts, true));
        ld[q].xt = reinterpret_cast<T *>(place(nullptr, n * ts, true));
        ld[q].res = reinterpret_cast<T *>(place(nullptr, n * ts, true));
    }
    if (!ok || !gpu::vcycle_tail_lds_fits<T>(static_cast<u32>(std::min<u64>(off, ~0u)), levels) ||
        off > gpu::kTailLdsMax)
        return;
    mg.tail_copies = std::move(plan);
    mg.tail_image = dbuf(std::max<u64>(image_bytes, 16));
    HIPCHK(hipMemcpy(mg.tail_image.p, ld.data(), ld.size() * sizeof(gpu::level_desc<T>),
                     hipMemcpyHostToDevice));
    mg.tail_image_bytes = static_cast<u32>(image_bytes);
    mg.tail_lds_bytes = static_cast<u32>(off);
    refresh_tail_image(mg);
}

template <typename T>
void build_levels(mk_sector_mg_s &mg, const std::vector<mk::sector_hierarchy> &H,
                  std::vector<level_dev<T>> &L) {
    const i64 K = static_cast<i64>(H.size());
    i64 depth = 0;
    for (const auto &h : H) depth = std::max<i64>(depth, static_cast<i64>(h.levels.size()));
    mg.max_levels = depth;
    // row offset of every active sector on every level
    std::vector<std::vector<i64>> offset(depth, std::vector<i64>(K, -1));
    std::vector<i64> level_rows(depth, 0);
    mg.sector_rows.assign(depth, std::vector<i64>(K, 0));
    for (i64 l = 0; l < depth; ++l)
        for (i64 k = 0; k < K; ++k)
            if (static_cast<i64>(H[k].levels.size()) > l) {
                offset[l][k] = level_rows[l];
                level_rows[l] += H[k].levels[l].rows;
                mg.sector_rows[l][k] = H[k].levels[l].rows;
            }
    L.resize(depth);
    // red-black: every level numbers each sector's states colour 0 first (ascending rank within
    // a colour); ord maps a level row to the sector's natural rank, pos the other way
    const bool by_colour = mg.opt.smoother == MK_RED_BLACK;
    std::vector<std::vector<std::vector<u32>>> ord(depth, std::vector<std::vector<u32>>(K)), pos = ord;
    if (by_colour) {
        for (i64 l = 0; l < depth; ++l)
            for (i64 k = 0; k < K; ++k) {
                if (offset[l][k] < 0) continue;
                const auto &c = H[k].levels[l].colour;
                if (static_cast<i64>(c.size()) != H[k].levels[l].rows)
                    throw std::logic_error("red-black smoother: missing colours");
                auto &o = ord[l][k];
                for (int want = 0; want < 2; ++want)
                    for (std::size_t i = 0; i < c.size(); ++i)
                        if (c[i] == want) o.push_back(static_cast<u32>(i));
                pos[l][k].resize(o.size());
                for (std::size_t i = 0; i < o.size(); ++i) pos[l][k][o[i]] = static_cast<u32>(i);
            }
    }
    auto nat = [&](i64 l, i64 k, i64 i) -> i64 { return by_colour ? ord[l][k][i] : i; };
    auto neu = [&](i64 l, i64 k, i64 i) -> i64 { return by_colour ? pos[l][k][i] : i; };
    for (i64 l = 0; l < depth; ++l) {
        auto &lv = L[l];
        lv.n = level_rows[l];
        if (lv.n >= (i64{1} << 32) - 1) throw std::overflow_error("multigrid level has too many states");
        // concatenated diffusion operator of the active sectors
        mk::generator_csr cat;
        cat.n = lv.n;
        cat.ptr.assign(1, 0);
        std::vector<u32> coarse_of;
        std::vector<double> weight;
        std::vector<std::uint8_t> colour;
        std::vector<u32> tstart, tsize;
        std::vector<u64> toff;
        std::vector<double> tinv;
        bool has_next = false;
        for (i64 k = 0; k < K; ++k) {
            if (offset[l][k] < 0) continue;
            const auto &s = H[k].levels[l];
            const u32 base = static_cast<u32>(offset[l][k]);
            lv.pre = s.pre_sweeps;
            lv.post = s.post_sweeps;
            // this sector's rows, in the level's order: lengths first, then a parallel fill
            const std::size_t r0 = cat.ptr.size() - 1;
            for (i64 i2 = 0; i2 < s.rows; ++i2) {
                const i64 i = nat(l, k, i2);
                cat.ptr.push_back(cat.ptr.back() + (s.A.ptr[i + 1] - s.A.ptr[i]));
            }
            cat.col.resize(cat.ptr.back());
            cat.val.resize(cat.ptr.back());
            cat.exit.resize(cat.ptr.size() - 1);
            if (!s.colour.empty()) colour.resize(cat.ptr.size() - 1);
            const i64 chunk = 1 << 15;
            mk::parallel_tasks((s.rows + chunk - 1) / chunk, [&](i64 t) {
                std::vector<std::pair<u32, double>> row;
                for (i64 i2 = t * chunk; i2 < std::min<i64>(s.rows, (t + 1) * chunk); ++i2) {
                    const i64 i = nat(l, k, i2);
                    row.clear();
                    for (u64 p = s.A.ptr[i]; p < s.A.ptr[i + 1]; ++p)
                        row.emplace_back(static_cast<u32>(base + neu(l, k, s.A.col[p])), s.A.val[p]);
                    if (by_colour) std::sort(row.begin(), row.end()); // columns ascending
                    u64 q = cat.ptr[r0 + i2];
                    for (const auto &[c, v] : row) {
                        cat.col[q] = c;
                        cat.val[q++] = v;
                    }
                    cat.exit[r0 + i2] = s.A.exit[i];
                    if (!s.colour.empty()) colour[r0 + i2] = s.colour[i];
                }
            });
            if (s.terminal) {
                const i64 m = s.rows;
                tstart.push_back(base);
                tsize.push_back(static_cast<u32>(m));
                toff.push_back(tinv.size());
                for (i64 i2 = 0; i2 < m; ++i2)
                    for (i64 j2 = 0; j2 < m; ++j2) tinv.push_back(s.inverse[nat(l, k, i2) * m + nat(l, k, j2)]);
                const std::size_t q0 = lv.h_tq.size();
                lv.h_tq.resize(q0 + static_cast<std::size_t>(m * m), 0.0);
                for (i64 i2 = 0; i2 < m; ++i2) {
                    const i64 i = nat(l, k, i2);
                    double *qrow = lv.h_tq.data() + q0 + i2 * m;
                    qrow[i2] += s.A.exit[i];
                    for (u64 p = s.A.ptr[i]; p < s.A.ptr[i + 1]; ++p) qrow[neu(l, k, s.A.col[p])] -= s.A.val[p];
                }
                lv.term_max = std::max<i64>(lv.term_max, m);
                coarse_of.insert(coarse_of.end(), m, 0xFFFFFFFFu);
                weight.insert(weight.end(), m, 0.0);
            } else {
                has_next = true;
                const i64 cb = offset[l + 1][k];
                for (i64 i2 = 0; i2 < s.rows; ++i2) {
                    const i64 i = nat(l, k, i2);
                    coarse_of.push_back(static_cast<u32>(cb + neu(l + 1, k, s.coarse_of[i])));
                    weight.push_back(s.weight[i]);
                }
                lv.n_fine_linked += s.rows;
            }
        }
        lv.A = upload_csr<T>(cat);
        if (mg.opt.smoother == MK_RED_BLACK) {
            if (static_cast<i64>(colour.size()) != lv.n) throw std::logic_error("red-black smoother: missing colours");
            split_colours<T>(cat, colour, lv);
            if (l == 0) {
                std::vector<u32> p0(lv.n);
                for (i64 k = 0; k < K; ++k) {
                    if (offset[0][k] < 0) continue;
                    for (i64 i2 = 0; i2 < H[k].levels[0].rows; ++i2)
                        p0[offset[0][k] + i2] = static_cast<u32>(offset[0][k] + nat(0, k, i2));
                }
                lv.perm = upload(p0);
                std::vector<u32> q0(lv.n);
                std::vector<T> e0(lv.n);
                for (i64 i = 0; i < lv.n; ++i) {
                    q0[p0[i]] = static_cast<u32>(i);
                    e0[p0[i]] = colour[i] == 0 ? static_cast<T>(cat.exit[i]) : T(-1);
                }
                lv.pos = upload(q0);
                lv.exit_box = upload(e0);
            }
        }
        if (mg.opt.smoother == MK_ILU0) build_ilu<T>(cat, mg.shift, lv);
        // aggregate-block smoother data of the sectors that are not terminal on this level
        i64 n_agg = 0;
        for (i64 k = 0; k < K; ++k)
            if (offset[l][k] >= 0 && H[k].levels[l].has_blocks)
                n_agg = std::max<i64>(n_agg, H[k].levels[l].blocks.aggregates);
        if (n_agg > 0) {
            std::vector<u32> members, member_block, block_first{0}, rows, jrows;
            std::vector<u64> inv_off;
            std::vector<T> inv;
            lv.h_agg_members.assign(1, 0);
            lv.h_agg_rows.assign(1, 0);
            for (i64 a = 0; a < n_agg; ++a) {
                for (i64 k = 0; k < K; ++k) {
                    if (offset[l][k] < 0 || !H[k].levels[l].has_blocks) continue;
                    const auto &bs = H[k].levels[l].blocks;
                    if (bs.aggregates != n_agg)
                        throw std::logic_error("sectors of one level disagree on the aggregates");
                    const u32 base = static_cast<u32>(offset[l][k]);
                    for (u64 b = bs.agg_blocks[a]; b < bs.agg_blocks[a + 1]; ++b) {
                        const u64 f = bs.block_first[b], e = bs.block_first[b + 1];
                        const u64 m = e - f;
                        const u32 bid = static_cast<u32>(block_first.size() - 1);
                        for (u64 g = f; g < e; ++g) {
                            members.push_back(base + bs.members[g]);
                            member_block.push_back(bid);
                        }
                        block_first.push_back(static_cast<u32>(members.size()));
                        inv_off.push_back(inv.size());
                        inv.insert(inv.end(), bs.inverse.begin() + bs.inv_offset[b],
                                   bs.inverse.begin() + bs.inv_offset[b] + m * m);
                    }
                    for (u64 q = bs.agg_rows[a]; q < bs.agg_rows[a + 1]; ++q) rows.push_back(base + bs.rows[q]);
                }
                lv.h_agg_members.push_back(members.size());
                lv.h_agg_rows.push_back(rows.size());
            }
            for (i64 k = 0; k < K; ++k) {
                if (offset[l][k] < 0 || H[k].levels[l].has_blocks) continue;
                for (i64 i = 0; i < H[k].levels[l].rows; ++i) jrows.push_back(static_cast<u32>(offset[l][k] + i));
            }
            if (members.size() >= (u64{1} << 32) || rows.size() >= (u64{1} << 32))
                throw std::overflow_error("aggregate-block smoother: too many block members");
            lv.n_agg = n_agg;
            lv.bs_members = upload(members);
            lv.bs_member_block = upload(member_block);
            lv.bs_block_first = upload(block_first);
            lv.bs_inv_off = upload(inv_off);
            lv.bs_inv = upload(inv);
            lv.bs_rows = upload(rows);
            lv.escr = dbuf(static_cast<std::size_t>(std::max<i64>(lv.n, 1)) * sizeof(T));
            HIPCHK(hipMemset(lv.escr.p, 0, lv.escr.bytes));
            lv.n_jrows = static_cast<i64>(jrows.size());
            if (lv.n_jrows > 0) lv.jrows = upload(jrows);
        }
        lv.n_term = static_cast<i64>(tstart.size());
        u64 term_rows = 0;
        for (u32 m : tsize) term_rows += m;
        lv.direct = l > 0 && !has_next && static_cast<i64>(term_rows) == lv.n;
        lv.h_tstart = tstart;
        lv.h_tsize = tsize;
        lv.h_toff = toff;
        if (lv.n_term > 0) {
            lv.tstart = upload(tstart);
            lv.tsize = upload(tsize);
            lv.toff = upload(toff);
            std::vector<T> inv_t(tinv.begin(), tinv.end());
            lv.tinv = upload(inv_t);
        }
        if (has_next) {
            const i64 nc = level_rows[l + 1];
            lv.n_next = nc;
            // children of every coarse row, fine rows ascending (the reference's scatter order)
            std::vector<u32> cptr(nc + 1, 0), cidx(lv.n_fine_linked);
            for (u32 c : coarse_of)
                if (c != 0xFFFFFFFFu) ++cptr[c + 1];
            for (i64 c = 0; c < nc; ++c) cptr[c + 1] += cptr[c];
            std::vector<u32> fillp(cptr.begin(), cptr.end() - 1);
            for (i64 i = 0; i < lv.n; ++i)
                if (coarse_of[i] != 0xFFFFFFFFu) cidx[fillp[coarse_of[i]]++] = static_cast<u32>(i);
            lv.coarse_of = upload(coarse_of);
            std::vector<T> w_t(weight.begin(), weight.end());
            lv.weight = upload(w_t);
            mk::generator_csr Rg;
            Rg.n = nc;
            Rg.ptr.assign(cptr.begin(), cptr.end());
            Rg.col = std::move(cidx);
            Rg.val.assign(Rg.col.size(), 1.0);
            Rg.exit.assign(nc, 0.0);
            lv.R = upload_csr<T>(Rg);
        }
        const std::size_t vb = static_cast<std::size_t>(std::max<i64>(lv.n, 1)) * sizeof(T);
        lv.rhs = dbuf(vb);
        lv.x = dbuf(vb);
        lv.xt = dbuf(vb);
        lv.res = dbuf(vb);
        if (mg.opt.coarse_cycles > 1) {
            lv.crhs = dbuf(vb);
            lv.cx = dbuf(vb);
        }
    }
    // the fused tail: the deepest run of levels that are small, Jacobi-smoothed and V-cycled
    std::vector<gpu::level_desc<T>> desc(depth);
    for (i64 l = 0; l < depth; ++l) {
        auto &lv = L[l];
        auto &d = desc[l];
        d.A = lv.A.view;
        d.R = lv.R.view;
        d.coarse_of = dp<u32>(lv.coarse_of);
        d.weight = dp<T>(lv.weight);
        d.n = lv.n;
        d.n_next = lv.n_next;
        d.pre = lv.pre;
        d.post = lv.post;
        d.n_term = lv.n_term;
        d.term_max = lv.term_max;
        d.tstart = dp<u32>(lv.tstart);
        d.tsize = dp<u32>(lv.tsize);
        d.toff = dp<u64>(lv.toff);
        d.tinv = dp<T>(lv.tinv);
        d.rhs = dp<T>(lv.rhs);
        d.x = dp<T>(lv.x);
        d.xt = dp<T>(lv.xt);
        d.res = dp<T>(lv.res);
        d.colour = lv.rb ? dp<std::uint8_t>(lv.colour) : nullptr;
        d.perm = dp<u32>(lv.perm);
        d.direct = lv.direct ? 1 : 0;
        for (int c = 0; c < 2 && lv.rb; ++c) {
            d.crows[c] = dp<u32>(lv.rowmap[c]);
            d.ncrows[c] = lv.Ac[c].n;
        }
    }
    mg.tail_start = -1;
    if (mg.opt.smoother == MK_JACOBI || mg.opt.smoother == MK_RED_BLACK ||
        mg.opt.smoother == MK_AGGREGATE_BLOCK) {
        // the tail's levels must each visit their coarse level once
        const i64 first_v = mg.opt.coarse_cycles == 1 ? 0 : cycle_levels(mg.opt);
        i64 start = depth;
        // (dense blocks of more than kWideTerminal rows take a wave per row: not in the tail)
        while (start > first_v && L[start - 1].n <= tail_rows() && L[start - 1].n_agg == 0 &&
               L[start - 1].term_max <= gpu::kWideTerminal)
            --start;
        // a lone direct level is one terminal-solve launch without the tail
        // (the tail tracks at most 32 levels)
        if (start < depth && depth - start <= 32 && !(start == depth - 1 && L[start].direct))
            mg.tail_start = start;
    }
    // a W visit from the level above the tail runs its coarse cycles inside the tail, which
    // accumulates the iterate in that level's cx
    mg.tail_w = mg.tail_start > 0 && mg.opt.coarse_cycles > 1 && mg.tail_start - 1 < cycle_levels(mg.opt);
    if (mg.tail_w) desc[mg.tail_start].acc = dp<T>(L[mg.tail_start].cx);
    mg.tail_desc = upload(desc);
    build_tail_lds<T>(mg, L, desc);
}

// aggregate-block smoothing (multigrid.hpp:309-342) of the level's block sectors: block
// Gauss-Seidel over the aggregates (in reverse for post-smoothing), block Jacobi within one
// aggregate, the residual kept in lv.res and corrected where each aggregate's update reaches;
// rows of sectors terminal on this level get Jacobi sweeps instead
template <typename T>
void block_smooth(const mk_sector_mg_s &mg, level_dev<T> &lv, std::size_t level, const T *rhs, T *x,
                  T *tmp, int sweeps, bool reverse, bool from_zero) {
    if (sweeps <= 0) return;
    auto &P = prof();
    auto cls = [&](kclass c) { return level == 0 ? c : K_COARSE; };
    const hipStream_t s = stream();
    const double shift = mg.shift, omega = mg.opt.jacobi_weight;
    const std::size_t ts = sizeof(T);
    T *r = dp<T>(lv.res);
    if (from_zero) {
        P.run(cls(K_OTHER), double(lv.n) * 2 * ts,
              [&] { HIPCHK(hipMemcpyAsync(r, rhs, lv.n * ts, hipMemcpyDeviceToDevice, s)); });
    } else {
        P.run(cls(K_RESIDUAL), op_bytes(lv.A, ts, ts, 2),
              [&] { gpu::level_residual<T>(lv.A.view, shift, rhs, x, r, s); });
    }
    const u32 *mem = dp<u32>(lv.bs_members), *rows = dp<u32>(lv.bs_rows);
    T *escr = dp<T>(lv.escr);
    for (int sweep = 0; sweep < sweeps; ++sweep) {
        for (i64 step = 0; step < lv.n_agg; ++step) {
            const i64 a = reverse ? lv.n_agg - 1 - step : step;
            const u64 m0 = lv.h_agg_members[a], m1 = lv.h_agg_members[a + 1];
            const u64 q0 = lv.h_agg_rows[a], q1 = lv.h_agg_rows[a + 1];
            P.run(cls(K_SWEEP), double(m1 - m0) * (3 * ts + 12), [&] {
                gpu::block_solve<T>(mem, dp<u32>(lv.bs_member_block), dp<u32>(lv.bs_block_first),
                                    dp<u64>(lv.bs_inv_off), dp<T>(lv.bs_inv), m0, m1, r, x, escr, s);
            });
            P.run(cls(K_SWEEP), double(q1 - q0) * (2 * ts + 4), [&] {
                gpu::block_residual<T>(lv.A.view, shift, rows, q0, q1, escr, r, s);
                gpu::block_zero<T>(mem, m0, m1, escr, s);
            });
        }
        if (lv.n_jrows > 0) {
            P.run(cls(K_SWEEP), double(lv.n_jrows) * 4 * ts, [&] {
                gpu::sweep_rows<T>(lv.A.view, shift, omega, rhs, x, tmp, dp<u32>(lv.jrows),
                                   static_cast<u64>(lv.n_jrows), s);
                gpu::copy_rows<T>(dp<u32>(lv.jrows), static_cast<u64>(lv.n_jrows), tmp, x, s);
            });
        }
    }
}

// one cycle of level l from a zero guess (population_multigrid::cycle, multigrid.hpp:201-228)
// on all active sectors at once. rhs is this level's right-hand side; rhs64 (level 0 only) is an
// FP64 input that the first sweep multiplies by scale and stores into lv.rhs in precision T. The
// result is written to out when given, else left in a level buffer whose pointer is returned.
// whether level l's first red-black half-sweep can ride on the kernel that produces its
// right-hand side: a red-black level with pre-smoothing, in natural order, run by the
// per-level kernels (not the fused tail, not a dense level)
template <typename T>
bool first_fusable(const mk_sector_mg_s &mg, const std::vector<level_dev<T>> &L, std::size_t l) {
    if (l >= L.size()) return false;
    const auto &lv = L[l];
    const bool tail = mg.tail_start >= 0 && static_cast<i64>(l) >= mg.tail_start && tail_switch().load() != 0;
    return lv.rb && lv.pre > 0 && !lv.direct && !lv.ilu && lv.n_agg == 0 && !tail && lv.perm.p == nullptr &&
           lv.colour.p != nullptr && fused_first_switch().load() != 0;
}

template <typename T>
T *run_level(const mk_sector_mg_s &mg, std::vector<level_dev<T>> &L, std::size_t l, const T *rhs,
             const double *rhs64, T *out, const double *scale = nullptr, int cycles = 1,
             bool first_done = false) {
    auto &lv = L[l];
    auto &P = prof();
    auto cls = [&](kclass c) { return l == 0 ? c : K_COARSE; };
    const hipStream_t s = stream();
    const double shift = mg.shift, omega = mg.opt.jacobi_weight;
    if (mg.tail_start >= 0 && static_cast<i64>(l) >= mg.tail_start && tail_switch().load() != 0) {
        // this level and every coarser one in one workgroup, from LDS where they fit
        const auto *desc = dp<gpu::level_desc<T>>(mg.tail_desc) + l;
        const bool lds = mg.tail_lds_bytes > 0 && static_cast<i64>(l) == mg.tail_start &&
                         tail_lds_switch().load() != 0;
        P.run(cls(K_SWEEP), 0.0, [&] {
            if (lds) {
                gpu::vcycle_tail_lds<T>(mg.tail_image.p, mg.tail_image_bytes, mg.tail_lds_bytes, static_cast<int>(L.size() - l),
                                        shift, omega, rhs, rhs64, scale, out, dp<T>(lv.x), s, cycles);
            } else {
                gpu::vcycle_tail<T>(desc, static_cast<int>(L.size() - l), shift, omega, rhs, rhs64,
                                    scale, out, s, cycles);
            }
        });
        return out ? out : dp<T>(lv.x);
    }
    if (cycles != 1) throw std::logic_error("sector_multigrid: coarse cycles inside a level outside the tail");
    if (lv.direct) {
        T *dst = out ? out : dp<T>(lv.x);
        P.run(cls(K_TERMINAL), 0.0, [&] {
            gpu::terminal_solve<T>(lv.n_term, dp<u32>(lv.tstart), dp<u32>(lv.tsize), dp<u64>(lv.toff),
                                   dp<T>(lv.tinv), lv.term_max, rhs, dst, s, true);
        });
        return dst;
    }
    const i64 n = lv.n;
    const std::size_t ts = sizeof(T);
    const double sweep_bytes = op_bytes(lv.A, ts, ts, 2);
    T *cur = dp<T>(lv.x);
    T *other = dp<T>(lv.xt);
    T *res = dp<T>(lv.res);
    const bool blocks = lv.n_agg > 0, rbgs = lv.rb, ilu = lv.ilu;
    const bool jacobi = !blocks && !rbgs && !ilu;
    // a red-black top level orders rows by colour: input and output are permuted
    const u32 *perm0 = dp<u32>(lv.perm);
    // the in-place smoothers work in the output when there is one (in the same row order)
    if ((rbgs || ilu) && out && !perm0) cur = out;
    bool have_residual = false;
    // one ILU(0) sweep: x += (L U)^-1 r, with u in `other`
    auto ilu_sweep = [&](const T *r) {
        const double fbytes = op_bytes(lv.A, ts, ts, 2) / 2.0 + double(n) * (4 + ts);
        P.run(cls(K_SWEEP), fbytes, [&] {
            for (const auto &g : lv.fw_launch)
                gpu::ilu_forward<T>(lv.A.view, dp<u32>(lv.fw_rows), dp<u32>(lv.fw_ptr), g.g0, g.g1,
                                    g.wide_rows, dp<T>(lv.pinv), r, other, s);
        });
        P.run(cls(K_SWEEP), fbytes + double(n) * 2 * ts, [&] {
            for (const auto &g : lv.bw_launch)
                gpu::ilu_backward<T>(lv.A.view, dp<u32>(lv.bw_rows), dp<u32>(lv.bw_ptr), g.g0, g.g1,
                                     g.wide_rows, dp<T>(lv.pinv), other, cur, s);
        });
    };
    auto half_bytes = [&](int c) { return op_bytes(lv.Ac[c], ts, ts, 2) + 4.0 * double(lv.Ac[c].n); };
    if (rbgs) {
        // red-black Gauss-Seidel: colour 0 from x = 0, then colour 1, colour 0, ...; after the last
        // colour-1 half-sweep only colour-0 rows have a residual
        if (lv.pre > 0) {
            T *rb = rhs64 || perm0 ? dp<T>(lv.rhs) : nullptr;
            // first_done: the kernel that produced rhs (restriction or residual) also ran this
            // first half-sweep into cur
            if (first_done && (rb || cycles != 1))
                throw std::logic_error("sector_multigrid: a fused first half-sweep on a level that reorders");
            // bytes: b, its row index or colour and exit code, the cast right-hand side, x of colour 0
            const double fb = rhs64 && perm0 ? double(n) * (8.0 + 4.0 + 2.5 * ts)
                                             : double(n) * ((rhs64 ? 8.0 + ts : double(ts)) + 2.0 + ts / 2.0 + (rb ? ts : 0));
            if (!first_done) P.run(cls(K_FIRST), fb, [&] {
                if (rhs64 && perm0) {
                    // from the caller's order: read b in its order and scatter to the rows
                    gpu::rb_first_scatter<T, double>(n, dp<T>(lv.exit_box), shift, rhs64, rb, cur, s, scale,
                                                     dp<u32>(lv.pos));
                } else if (rhs64) {
                    gpu::rb_first<T, double>(lv.A.view, dp<std::uint8_t>(lv.colour), shift, rhs64, rb, cur, s,
                                             scale, perm0);
                } else {
                    gpu::rb_first<T, T>(lv.A.view, dp<std::uint8_t>(lv.colour), shift, rhs, rb, cur, s,
                                        nullptr, perm0);
                }
            });
            if (rb) rhs = rb;
            for (int half = 1; half < 2 * lv.pre; ++half) {
                const int c = half & 1;
                T *rz = half == 2 * lv.pre - 1 ? res : nullptr;
                P.run(cls(K_SWEEP), half_bytes(c), [&] { gpu::gs_half<T>(lv.Ac[c].view, shift, rhs, cur, rz, s); });
            }
            P.run(cls(K_RESIDUAL), half_bytes(0), [&] {
                gpu::level_residual<T>(lv.Ac[0].view, shift, rhs, cur, res, s);
            });
            have_residual = true;
        } else {
            if (rhs64 || perm0) {
                T *rb = dp<T>(lv.rhs);
                P.run(cls(K_OTHER), double(n) * (8 + ts), [&] {
                    if (perm0 && rhs64) {
                        gpu::permute_in<T, double>(n, perm0, rhs64, rb, s, scale);
                    } else if (perm0) {
                        gpu::permute_in<T, T>(n, perm0, rhs, rb, s);
                    } else {
                        gpu::convert<double, T>(n, rhs64, rb, s, scale);
                    }
                });
                rhs = rb;
            }
            P.run(cls(K_OTHER), double(n) * ts, [&] { gpu::fill<T>(n, T(0), cur, s); });
        }
    } else if (ilu) {
        if (rhs64) {
            T *rb = dp<T>(lv.rhs);
            P.run(cls(K_OTHER), double(n) * (8 + ts), [&] { gpu::convert<double, T>(n, rhs64, rb, s, scale); });
            rhs = rb;
        }
        P.run(cls(K_OTHER), double(n) * ts, [&] { gpu::fill<T>(n, T(0), cur, s); });
        for (int k = 0; k < lv.pre; ++k) {
            if (k > 0) {
                P.run(cls(K_RESIDUAL), sweep_bytes, [&] { gpu::level_residual<T>(lv.A.view, shift, rhs, cur, res, s); });
            }
            ilu_sweep(k == 0 ? rhs : res); // from x = 0 the residual is rhs
        }
    } else if (blocks) {
        if (rhs64) {
            T *rb = dp<T>(lv.rhs);
            P.run(cls(K_OTHER), double(n) * (8 + ts),
                  [&] { gpu::convert<double, T>(n, rhs64, rb, s, scale); });
            rhs = rb;
        }
        P.run(cls(K_OTHER), double(n) * ts, [&] { gpu::fill<T>(n, T(0), cur, s); });
        block_smooth<T>(mg, lv, l, rhs, cur, other, lv.pre, false, true);
    } else if (rhs64) {
        T *rb = dp<T>(lv.rhs);
        if (lv.pre > 0) {
            P.run(cls(K_FIRST), double(n) * (8 + 3 * ts), [&] {
                gpu::jacobi_first<T, double>(lv.A.view, shift, omega, rhs64, rb, cur, s, scale);
            });
        } else {
            P.run(cls(K_OTHER), double(n) * (8 + ts),
                  [&] { gpu::convert<double, T>(n, rhs64, rb, s, scale); });
        }
        rhs = rb;
    } else if (lv.pre > 0) {
        P.run(cls(K_FIRST), double(n) * 3 * ts, [&] {
            gpu::jacobi_first<T, T>(lv.A.view, shift, omega, rhs, nullptr, cur, s);
        });
    }
    if (jacobi && lv.pre == 0) P.run(cls(K_OTHER), double(n) * ts, [&] { gpu::fill<T>(n, T(0), cur, s); });
    for (int k = 1; jacobi && k < lv.pre; ++k) {
        P.run(cls(K_SWEEP), sweep_bytes, [&] {
            gpu::jacobi_sweep<T>(lv.A.view, shift, omega, rhs, cur, other, s);
        });
        std::swap(cur, other);
    }
    if (!have_residual) {
        P.run(cls(K_RESIDUAL), sweep_bytes, [&] {
            gpu::level_residual<T>(lv.A.view, shift, rhs, cur, res, s);
        });
    }
    if (lv.n_term > 0) {
        P.run(cls(K_TERMINAL), 0.0, [&] {
            gpu::terminal_solve<T>(lv.n_term, dp<u32>(lv.tstart), dp<u32>(lv.tsize),
                                   dp<u64>(lv.toff), dp<T>(lv.tinv), lv.term_max, res, cur, s);
        });
    }
    if (lv.n_next > 0) {
        auto &c = L[l + 1];
        // restriction reads each linked fine row once (x) plus R's encoding, writes the coarse rows
        const double rbytes = op_bytes(lv.R, 0, ts, 1) + double(lv.n_fine_linked) * ts;
        // red-black with post-smoothing: only colour-0 rows need the correction, since the first
        // post-smoothing half-sweep rewrites colour-1 rows without reading them
        const bool red_only = rbgs && lv.post > 0;
        const i64 np = red_only ? lv.Ac[0].n : n;
        const u32 *prows = red_only ? dp<u32>(lv.rowmap[0]) : nullptr;
        const double pbytes = double(np) * (4 + 3 * ts + (red_only ? 4 : 0)) + double(c.n) * ts;
        if (mg.opt.coarse_cycles <= 1 || l >= cycle_levels(mg.opt)) {
            // the coarse level's first red-black half-sweep rides on the restriction
            const bool fuse = first_fusable(mg, L, l + 1);
            P.run(cls(K_RESTRICT), rbytes, [&] {
                if (fuse) {
                    gpu::gather_sum_first<T>(lv.R.view, res, dp<T>(c.rhs), c.A.view,
                                             dp<std::uint8_t>(c.colour), shift, dp<T>(c.x), s);
                } else {
                    gpu::gather_sum<T>(lv.R.view, res, dp<T>(c.rhs), s);
                }
            });
            // the coarse levels' launches are timed as one interval per visit (an event pair per
            // small launch would cost as much as the launch)
            T *cx = nullptr;
            // a direct coarse level in wave-per-row blocks: its dense solve prolongs each coarse
            // value to the colour-0 children right away (one launch instead of two)
            const bool dense_fuse = red_only && c.direct && c.term_max > gpu::kWideTerminal &&
                                    lv.colour.p != nullptr && coarse_fuse_switch().load() != 0;
            if (dense_fuse) {
                P.span(K_COARSE, [&] {
                    gpu::terminal_prolong<T>(c.n_term, dp<u32>(c.tstart), dp<u32>(c.tsize), dp<u64>(c.toff),
                                             dp<T>(c.tinv), c.term_max, dp<T>(c.rhs), dp<T>(c.x), lv.R.view,
                                             dp<std::uint8_t>(lv.colour), dp<T>(lv.weight), cur, s);
                });
            } else {
                P.span(K_COARSE, [&] {
                    cx = run_level<T>(mg, L, l + 1, dp<T>(c.rhs), nullptr, nullptr, nullptr, 1, fuse);
                });
                P.run(cls(K_PROLONG), pbytes, [&] {
                    gpu::prolong_add<T>(np, dp<u32>(lv.coarse_of), dp<T>(lv.weight), cx, cur, s, prows);
                });
            }
        } else {
            // the tail runs all of the coarse cycles in one launch
            const bool in_tail = mg.tail_w && mg.tail_start == static_cast<i64>(l + 1) &&
                                 tail_switch().load() != 0;
            // the first coarse cycle writes cx directly, so its first half-sweep, fused into the
            // restriction, goes there too
            const bool fuse = !in_tail && first_fusable(mg, L, l + 1);
            P.run(cls(K_RESTRICT), rbytes, [&] {
                if (fuse) {
                    gpu::gather_sum_first<T>(lv.R.view, res, dp<T>(c.crhs), c.A.view,
                                             dp<std::uint8_t>(c.colour), shift, dp<T>(c.cx), s);
                } else {
                    gpu::gather_sum<T>(lv.R.view, res, dp<T>(c.crhs), s);
                }
            });
            // from cx = 0 the first coarse residual is crhs itself and the first correction
            // is cx: that cycle writes cx directly
            const double cbytes = op_bytes(c.A, ts, ts, 2);
            // the last cycle's correction is added inside the prolongation (cx + ce, then the
            // weighted add, as the separate axpy and prolongation round)
            T *last = nullptr;
            // every launch of the coarse cycles works on the coarse level: one timed interval
            P.span(K_COARSE, [&] {
                run_level<T>(mg, L, l + 1, dp<T>(c.crhs), nullptr, dp<T>(c.cx), nullptr,
                             in_tail ? mg.opt.coarse_cycles : 1, fuse);
                for (int cyc = 1; !in_tail && cyc < mg.opt.coarse_cycles; ++cyc) {
                    // the next cycle's first half-sweep rides on its residual (into c.x)
                    P.run(K_COARSE, cbytes, [&] {
                        if (fuse) {
                            gpu::level_residual_first<T>(c.A.view, shift, dp<T>(c.crhs), dp<T>(c.cx),
                                                         dp<T>(c.rhs), c.A.view,
                                                         dp<std::uint8_t>(c.colour), dp<T>(c.x), s);
                        } else {
                            gpu::level_residual<T>(c.A.view, shift, dp<T>(c.crhs), dp<T>(c.cx),
                                                   dp<T>(c.rhs), s);
                        }
                    });
                    T *ce = run_level<T>(mg, L, l + 1, dp<T>(c.rhs), nullptr, nullptr, nullptr, 1, fuse);
                    if (cyc + 1 < mg.opt.coarse_cycles) {
                        P.run(K_COARSE, double(c.n) * 3 * ts, [&] { gpu::axpy<T>(c.n, T(1), ce, dp<T>(c.cx), s); });
                    } else {
                        last = ce;
                    }
                }
            });
            P.run(cls(K_PROLONG), pbytes, [&] {
                gpu::prolong_add<T>(np, dp<u32>(lv.coarse_of), dp<T>(lv.weight), dp<T>(c.cx),
                                    cur, s, prows, last);
            });
        }
    }
    if (blocks) block_smooth<T>(mg, lv, l, rhs, cur, other, lv.post, true, false);
    for (int half = 0; rbgs && half < 2 * lv.post; ++half) {
        const int c = 1 - (half & 1); // reverse colour order
        // the last two half-sweeps also store their rows to the output in the caller's order
        T *nat = perm0 && out && half >= 2 * lv.post - 2 ? out : nullptr;
        P.run(cls(K_SWEEP), half_bytes(c), [&] {
            gpu::gs_half<T>(lv.Ac[c].view, shift, rhs, cur, nullptr, s, nat, perm0);
        });
    }
    if (perm0 && out) {
        if (lv.post == 0) P.run(cls(K_OTHER), double(n) * (2 * ts + 4), [&] { gpu::permute_out<T>(n, perm0, cur, out, s); });
        return out;
    }
    for (int k = 0; ilu && k < lv.post; ++k) {
        P.run(cls(K_RESIDUAL), sweep_bytes, [&] { gpu::level_residual<T>(lv.A.view, shift, rhs, cur, res, s); });
        ilu_sweep(res);
    }
    for (int k = 0; jacobi && k < lv.post; ++k) {
        T *dst = (k == lv.post - 1 && out) ? out : other;
        P.run(cls(K_SWEEP), sweep_bytes, [&] {
            gpu::jacobi_sweep<T>(lv.A.view, shift, omega, rhs, cur, dst, s);
        });
        if (dst == other) {
            std::swap(cur, other);
        } else {
            cur = dst;
        }
    }
    if (out && cur != out) {
        P.run(cls(K_OTHER), double(n) * 2 * ts,
              [&] { HIPCHK(hipMemcpyAsync(out, cur, n * ts, hipMemcpyDeviceToDevice, s)); });
        cur = out;
    }
    return cur;
}

// z = M (scale b) on the device: b FP64, z in the hierarchy's precision; scale is read from
// device memory (1 when null)
void mg_apply_device(mk_sector_mg_s &mg, const double *b, void *z, const double *scale = nullptr) {
    if (mg.precision == MK_FP32) {
        run_level<float>(mg, mg.L32, 0, nullptr, b, static_cast<float *>(z), scale);
    } else if (!scale) {
        run_level<double>(mg, mg.L64, 0, b, nullptr, static_cast<double *>(z));
    } else {
        run_level<double>(mg, mg.L64, 0, nullptr, b, static_cast<double *>(z), scale);
    }
}

// ------------------------------------------------------------------ flexible GMRES
struct gmres_workspace {
    i64 n = 0;
    i64 ld = 0; // basis stride: n rounded up to 64 elements, so every vector starts on a 512 B boundary
    i64 m = 0;
    std::size_t zs = 0;
    dbuf V, Z, w, partial, hd, yd;
    dbuf rpart; // one partial residual norm per operator row block
    i64 rpart_n = 0;
    dbuf scal;                 // per-iteration scalars read by the kernels: 1 / beta_j, beta_j
    double *hscal = nullptr;   // their pinned host copy
    graph_cache graphs;        // captured iterations (pointers into this workspace)
    // device scalars hd and their pinned host copy: pass-1 projections at 0, pass-2
    // projections and |w|^2 at m + 2, (beta, flag) at 2 (m + 2); y follows at 3 (m + 2)
    double *host = nullptr;
    ~gmres_workspace() {
        graphs.clear();
        if (host) (void)hipHostFree(host);
        if (hscal) (void)hipHostFree(hscal);
    }
    void ensure(i64 n_, i64 m_, std::size_t zs_) {
        if (n_ <= n && m_ <= m && (zs_ == 0 || zs_ <= zs) && V.p) return;
        n = std::max(n, n_);
        m = std::max(m, m_);
        zs = std::max(zs, zs_);
        V = dbuf();
        Z = dbuf();
        ld = (n + 63) / 64 * 64;
        V = dbuf(static_cast<std::size_t>(ld) * (m + 1) * sizeof(double));
        if (zs) Z = dbuf(static_cast<std::size_t>(ld) * m * zs);
        w = dbuf(n * sizeof(double));
        const int nb = gpu::reduction_blocks(n);
        partial = dbuf(static_cast<std::size_t>(nb) * (m + 2) * sizeof(double));
        hd = dbuf(3 * (m + 2) * sizeof(double));
        yd = dbuf((m + 1) * sizeof(double));
        if (host) (void)hipHostFree(host);
        host = nullptr;
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&host), (4 * (m + 2)) * sizeof(double)));
        if (!hscal) HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&hscal), 2 * sizeof(double)));
        scal = dbuf(2 * sizeof(double));
        graphs.clear();
    }
};

// one Krylov workspace per host thread
gmres_workspace &workspace() {
    thread_local gmres_workspace w;
    return w;
}

// The device and host arrays one FGMRES call works in: the thread's cached workspace (pinned
// host scalars, graph replay) or a view into the allocation-free scratch array (pageable host
// scalars on the caller's stack, per-iteration scalars stored by kernel arguments, eager launches)
struct krylov_space {
    i64 ld = 0;                // basis stride (elements)
    double *V = nullptr;       // m + 1 basis vectors
    void *Z = nullptr;         // m preconditioned vectors
    double *w = nullptr, *partial = nullptr, *hd = nullptr, *yd = nullptr, *rpart = nullptr,
           *scal = nullptr;
    double *host = nullptr;    // 4 (m + 2) doubles of host memory the device scalars are read into
    double *hscal = nullptr;   // 2 pinned doubles the per-iteration scalars are copied from; null:
                               // they are stored by a kernel's arguments
    graph_cache *graphs = nullptr; // null: eager launches only
};

krylov_space cached_space(gmres_workspace &ws) {
    krylov_space K;
    K.ld = ws.ld;
    K.V = dp<double>(ws.V);
    K.Z = ws.Z.p;
    K.w = dp<double>(ws.w);
    K.partial = dp<double>(ws.partial);
    K.hd = dp<double>(ws.hd);
    K.yd = dp<double>(ws.yd);
    K.rpart = dp<double>(ws.rpart);
    K.scal = dp<double>(ws.scal);
    K.host = ws.host;
    K.hscal = ws.hscal;
    K.graphs = &ws.graphs;
    return K;
}

// up to kFusedBasis doubles passed by value into device memory (no host staging buffer)
struct value_pack {
    double v[gpu::kFusedBasis];
};

__global__ void k_store_values(value_pack p, int count, double *dst) {
    const int i = static_cast<int>(threadIdx.x);
    if (i < count) dst[i] = p.v[i];
}

void store_values(const double *vals, int count, double *dst, hipStream_t s) {
    if (count <= 0) return;
    if (count > gpu::kFusedBasis) throw std::logic_error("store_values: too many values");
    value_pack p{};
    std::copy(vals, vals + count, p.v);
    hipLaunchKernelGGL(k_store_values, dim3(1), dim3(gpu::kFusedBasis), 0, s, p, count, dst);
    check_launch("store_values");
}

double device_norm2(const double *partial, int nb, double *hd) {
    auto &P = prof();
    P.run(K_DOT, nb * 8.0, [&] { gpu::reduce_partials(partial, nb, 1, hd, stream()); });
    double v = 0.0;
    HIPCHK(hipMemcpyAsync(&v, hd, sizeof(double), hipMemcpyDeviceToHost, stream()));
    drain();
    return v;
}

// Flexible GMRES(m), right-preconditioned: x = x0 + Z y with z_j = M v_j (the numerics gmres
// with a preconditioner, called at tests/test_rdme.cpp:242-243). The Arnoldi basis is
// orthogonalised by CGS2 or, for preconditioned solves with m <= kFusedBasis, by delayed CGS2:
// step j applies M and A to the candidate u_j = Q s_j + beta_j q_j (orthogonalised once, scaled
// by 1 / beta_j), and one reduction gives Q^T w and u_j^T w; the update kernel then finalises
// q_j and forms the next candidate with its projections. Because w_j = A z_j is expanded as
// Q (t + s_{j+1}) + beta_{j+1} q_{j+1}, column j of H is (t + s_{j+1}, beta_{j+1}) as in CGS2.
i64 gmres_restart(const mk_gmres_options &o, i64 n) {
    return std::max<i64>(1, std::min<i64>(o.restart, n));
}

// whether the whole solve runs in one workgroup: a sector hierarchy of precision ZT whose levels,
// the top one included, sit in the LDS tail, one row per thread and a restart the kernel holds
template <typename ZT>
bool tiny_fits(const precond *M, i64 n, i64 m) {
    if (!M || !M->mg || tiny_gmres_switch().load() == 0) return false;
    const mk_sector_mg_s &mg = *M->mg;
    const bool precision = (mg.precision == MK_FP32) == std::is_same_v<ZT, float>;
    return precision && mg.tail_start == 0 && mg.tail_lds_bytes > 0 && tail_switch().load() != 0 &&
           tail_lds_switch().load() != 0 && n <= gpu::kTinyRows && m <= gpu::kTinyBasis &&
           gpu::gmres_tiny_lds(mg.tail_lds_bytes) > 0;
}

// restarted FGMRES (CGS2, no delayed variant) with the LDS tail, in one kernel launch
template <typename ZT>
mk_solver_result fgmres_tiny(mk_generator_s &A, double shift, const mk_sector_mg_s &mg, const double *b,
                             double *x, const mk_gmres_options &o, const krylov_space &K, i64 m) {
    gpu::tiny_gmres_args a;
    a.A = A.A.view;
    a.shift = shift;
    a.b = b;
    a.x = x;
    a.V = K.V;
    a.Z = K.Z;
    a.ld = K.ld;
    a.n = static_cast<int>(A.n);
    a.m = static_cast<int>(m);
    a.max_iterations = static_cast<int>(std::min<i64>(o.max_iterations, i64{1} << 30));
    a.tolerance = o.tolerance;
    a.result = K.hd;
    a.image = mg.tail_image.p;
    a.image_bytes = mg.tail_image_bytes;
    a.lds_bytes = mg.tail_lds_bytes;
    a.levels = static_cast<int>(mg.precision == MK_FP32 ? mg.L32.size() : mg.L64.size());
    a.omega = mg.opt.jacobi_weight;
    a.n_ent = A.A.padded;
    const hipStream_t s = stream();
    prof().run(K_OTHER, 0.0, [&] { gpu::gmres_tiny<ZT>(a, s); });
    HIPCHK(hipMemcpyAsync(K.host, K.hd, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    drain();
    mk_solver_result res{static_cast<int64_t>(K.host[0]), K.host[1], static_cast<int32_t>(K.host[2])};
    return res;
}

template <typename ZT>
mk_solver_result fgmres(mk_generator_s &A, double shift, const precond *M, const double *b,
                        double *x, const mk_gmres_options &o, const krylov_space &K) {
    const i64 n = A.n;
    const i64 m = gmres_restart(o, n);
    require(m + 1 <= 512, "gmres: restart above 511 is not supported on the device");
    require(o.orthogonalization == MK_DCGS2 || o.orthogonalization == MK_CGS2,
            "gmres: unknown orthogonalization");
    if (o.max_iterations > 0 && tiny_fits<ZT>(M, n, m)) return fgmres_tiny<ZT>(A, shift, *M->mg, b, x, o, K, m);
    const bool delayed = M && o.orthogonalization == MK_DCGS2 && m <= gpu::kFusedBasis;
    auto &P = prof();
    const hipStream_t s = stream();
    const int nb = gpu::reduction_blocks(n);
    double *V = K.V, *w = K.w;
    double *partial = K.partial, *hd = K.hd;
    ZT *Z = M ? static_cast<ZT *>(K.Z) : nullptr;
    const i64 ld = K.ld;
    const double spmv_bytes = op_bytes(A.A, 8, 8, 1);
    const i64 hs = m + 2; // stride of the device scalar slots in hd

    mk_solver_result res{0, 0.0, 0};
    // V_0 = r = b - A x and ||r||^2 (left in hd[0])
    const int rb = static_cast<int>(std::max<i64>(A.A.view.n_blocks, 1));
    auto true_residual = [&]() {
        P.run(K_SPMV, spmv_bytes + n * 8.0, [&] {
            gpu::shifted_apply<double>(A.A.view, shift, x, b, V, K.rpart, s);
        });
        return std::sqrt(device_norm2(K.rpart, rb, hd));
    };
    // w = A M (scale v), scale read from device memory (1 when null)
    auto operator_image = [&](i64 j, const double *v, const double *scale) {
        if (M) {
            ZT *Zj = Z + j * ld;
            M->apply(v, Zj, scale);
            P.run(K_SPMV, op_bytes(A.A, sizeof(ZT), 8, 1), [&] {
                gpu::shifted_apply<ZT>(A.A.view, shift, Zj, nullptr, w, nullptr, s);
            });
        } else {
            P.run(K_SPMV, spmv_bytes, [&] {
                gpu::shifted_apply<double>(A.A.view, shift, v, nullptr, w, nullptr, s);
            });
        }
    };
    auto fetch = [&](std::size_t count) {
        HIPCHK(hipMemcpyAsync(K.host, hd, count * sizeof(double), hipMemcpyDeviceToHost, s));
        drain();
    };
    // device copies of host values the next kernels read: from the pinned staging array, or
    // through a kernel's arguments
    auto put = [&](const double *vals, int count, double *dst, double *pinned) {
        if (pinned) {
            std::copy(vals, vals + count, pinned);
            HIPCHK(hipMemcpyAsync(dst, pinned, count * sizeof(double), hipMemcpyHostToDevice, s));
        } else {
            store_values(vals, count, dst, s);
        }
    };
    // the body of one Arnoldi step: replayed from a captured graph, or launched eagerly
    auto step = [&](const graph_key &key, auto &&body) {
        if (K.graphs) {
            K.graphs->run(key, s, body, n);
        } else {
            body();
        }
    };
    double beta = true_residual();
    res.residual = beta;
    // converged already, or a non-finite right-hand side or guess: nothing to iterate on
    if (!(beta > o.tolerance) || !std::isfinite(beta)) {
        res.converged = std::isfinite(beta) ? 1 : 0;
        return res;
    }
    hessenberg_ls ls(m);
    std::vector<double> col(m + 1);
    while (res.iterations < o.max_iterations) {
        ls.start(beta);
        i64 j = 0;
        if (delayed) {
            // V_0 holds the candidate u_0 = r with s_0 empty and beta_0 = |r|
            double bj = beta;
            double *sd = hd + hs, *sd_next = hd + 2 * hs;
            for (; j < m && res.iterations < o.max_iterations; ++j) {
                ++res.iterations;
                const int cnt = static_cast<int>(j);
                const double sv[2] = {1.0 / bj, bj};
                double *scal = K.scal;
                put(sv, 2, scal, K.hscal);
                // the fused-tail switches choose the launches captured: both are part of the key
                graph_key key{A.uid, M ? M->uid : 0, j,
                              static_cast<i64>(sd > sd_next) + 2 * tail_switch().load() +
                                  4 * tail_lds_switch().load() + 8 * fused_first_switch().load() +
                                  16 * coarse_fuse_switch().load(),
                              n, m, bits(shift)};
                step(key, [&] {
                    operator_image(j, V + j * ld, scal);
                    // t = Q^T w and u_j^T w in hd[0 .. j]
                    P.run(K_DOT, double(cnt + 2) * n * 8.0, [&] {
                        gpu::multi_dot(V, ld, cnt + 1, w, n, partial, nb, s);
                        gpu::reduce_partials(partial, nb, cnt + 1, hd, s);
                    });
                    // q_j, eta (hd[m + 1]), u_{j+1} and its projections s_{j+1}, |u_{j+1}|^2
                    P.run(K_UPDATE, double(cnt + 4) * n * 8.0, [&] {
                        gpu::dcgs_update(V, ld, cnt, sd, hd, scal + 1, w, n, partial, nb, hd + m + 1, s);
                        gpu::reduce_partials(partial, nb, cnt + 2, sd_next, s);
                    });
                });
                fetch(3 * hs);
                const double *t = K.host, *sn = K.host + (sd_next - hd);
                const double eta = K.host[m + 1];
                for (i64 k = 0; k < j; ++k) col[k] = t[k] + sn[k];
                col[j] = eta + sn[j];
                double ss = 0.0;
                for (i64 k = 0; k <= j; ++k) ss += sn[k] * sn[k];
                double alpha = sn[j + 1];
                if (!(ss <= 0.25 * alpha)) {
                    // the candidate kept too much of span(Q) for the Pythagorean norm: project
                    // it once more explicitly and keep the remaining coefficients for the next
                    // step's finalisation
                    double *u = V + (j + 1) * ld;
                    P.run(K_UPDATE, double(cnt + 3) * n * 8.0, [&] {
                        gpu::basis_update(u, V, ld, cnt + 1, sd_next, n, nullptr, nb, s);
                    });
                    P.run(K_DOT, double(cnt + 3) * n * 8.0, [&] {
                        gpu::multi_dot(V, ld, cnt + 2, u, n, partial, nb, s);
                        gpu::reduce_partials(partial, nb, cnt + 2, sd_next, s);
                    });
                    fetch(3 * hs);
                    ss = 0.0;
                    for (i64 k = 0; k <= j; ++k) {
                        col[k] += sn[k];
                        ss += sn[k] * sn[k];
                    }
                    alpha = sn[j + 1];
                }
                const double hn = std::sqrt(std::max(alpha - ss, 0.0));
                std::swap(sd, sd_next);
                bj = hn;
                const double est = ls.add_column(j, col.data(), hn);
                if (!(est > o.tolerance) || !(hn > 0.0) || !std::isfinite(hn)) {
                    ++j;
                    break;
                }
            }
        } else {
            // v_0 = r / ||r||, with ||r||^2 still in hd[0] from the true residual
            P.run(K_OTHER, n * 16.0, [&] { gpu::scale_by_norm(n, V, hd, V, s); });
            for (; j < m && res.iterations < o.max_iterations; ++j) {
                ++res.iterations;
                double *Vj = V + j * ld;
                operator_image(j, Vj, nullptr);
                // classical Gram-Schmidt, twice (CGS2)
                const int cnt = static_cast<int>(j + 1);
                double *h1d = hd, *cwd = hd + hs, *bfd = hd + 2 * hs;
                double *vnext = V + (j + 1) * ld;
                const bool fused = cnt <= gpu::kFusedBasis;
                if (fused) {
                    // pass 1 projections; update fused with the pass-2 projections and |w|^2;
                    // pass-2 update fused with the normalisation
                    P.run(K_DOT, double(cnt + 1) * n * 8.0, [&] {
                        gpu::multi_dot(V, ld, cnt, w, n, partial, nb, s);
                        gpu::reduce_partials(partial, nb, cnt, h1d, s);
                    });
                    P.run(K_UPDATE, double(cnt + 2) * n * 8.0, [&] {
                        gpu::update_dot(w, V, ld, cnt, h1d, n, partial, nb, s);
                        gpu::reduce_partials(partial, nb, cnt + 1, cwd, s);
                    });
                    P.run(K_UPDATE, double(cnt + 2) * n * 8.0, [&] {
                        gpu::update_scale(w, V, ld, cnt, cwd, n, vnext, bfd, s);
                    });
                } else {
                    for (int pass = 0; pass < 2; ++pass) {
                        double *h = pass == 0 ? h1d : cwd;
                        P.run(K_DOT, double(cnt + 1) * n * 8.0, [&] {
                            gpu::multi_dot(V, ld, cnt, w, n, partial, nb, s);
                            gpu::reduce_partials(partial, nb, cnt, h, s);
                        });
                        double *np = pass == 1 ? partial + static_cast<i64>(nb) * (m + 1) : nullptr;
                        P.run(K_UPDATE, double(cnt + 2) * n * 8.0, [&] {
                            gpu::basis_update(w, V, ld, cnt, h, n, np, nb, s);
                        });
                    }
                    P.run(K_DOT, nb * 8.0, [&] {
                        gpu::reduce_partials(partial + static_cast<i64>(nb) * (m + 1), nb, 1,
                                             bfd + 2, s);
                    });
                    P.run(K_OTHER, n * 16.0, [&] { gpu::scale_by_norm(n, w, bfd + 2, vnext, s); });
                }
                fetch(3 * hs);
                const double *h1 = K.host, *h2 = K.host + hs;
                double hn;
                if (fused) {
                    hn = K.host[2 * hs];
                    if (K.host[2 * hs + 1] != 0.0) {
                        // the second pass removed too much for the Pythagorean norm: v_{j+1}
                        // holds the unnormalised vector; normalise it explicitly
                        P.run(K_DOT, n * 8.0, [&] {
                            gpu::multi_dot(vnext, n, 1, vnext, n, partial, nb, s);
                            gpu::reduce_partials(partial, nb, 1, bfd + 2, s);
                        });
                        P.run(K_OTHER, n * 16.0,
                              [&] { gpu::scale_by_norm(n, vnext, bfd + 2, vnext, s); });
                        fetch(3 * hs);
                        hn = std::sqrt(K.host[2 * hs + 2]);
                    }
                } else {
                    hn = std::sqrt(K.host[2 * hs + 2]);
                }
                for (i64 k = 0; k <= j; ++k) col[k] = h1[k] + h2[k];
                const double est = ls.add_column(j, col.data(), hn);
                if (!(est > o.tolerance) || !(hn > 0.0) || !std::isfinite(hn)) {
                    ++j;
                    break;
                }
            }
        }
        ls.solve(j);
        for (i64 c = 0; c < j; c += gpu::kFusedBasis) {
            const int cnt = static_cast<int>(std::min<i64>(gpu::kFusedBasis, j - c));
            put(ls.y.data() + c, cnt, K.yd + c, K.hscal ? K.host + 3 * hs + c : nullptr);
        }
        P.run(K_OTHER, double(j) * n * (M ? sizeof(ZT) : 8) + n * 16.0, [&] {
            if (M) {
                gpu::combine<ZT>(x, Z, ld, static_cast<int>(j), K.yd, n, s);
            } else {
                gpu::combine<double>(x, V, ld, static_cast<int>(j), K.yd, n, s);
            }
        });
        beta = true_residual();
        res.residual = beta;
        if (!std::isfinite(beta)) break;
        if (!(beta > o.tolerance)) {
            res.converged = 1;
            break;
        }
    }
    return res;
}

precond sector_precond(mk_sector_mg_s *mg) {
    return precond{mg->uid, mg->rows, mg->shift, mg->precision,
                   [mg](const double *v, void *z, const double *scale) { mg_apply_device(*mg, v, z, scale); },
                   mg};
}

mk_solver_result gmres_device(mk_generator_s &A, double shift, mk_sector_mg_s *mg, const double *b,
                              double *x, const mk_gmres_options &o) {
    if (!mg) return mk::rt::gmres_device(A, shift, static_cast<const precond *>(nullptr), b, x, o);
    const precond M = sector_precond(mg);
    return mk::rt::gmres_device(A, shift, &M, b, x, o);
}

} // namespace

namespace {
// checks shared by both workspaces; the size of a preconditioned vector (0 without one)
std::size_t check_gmres(const mk_generator_s &A, double shift, const precond *M,
                        const mk_gmres_options &o) {
    require(gmres_restart(o, A.n) + 1 <= 512, "gmres: restart above 511 is not supported on the device");
    require(o.orthogonalization == MK_DCGS2 || o.orthogonalization == MK_CGS2,
            "gmres: unknown orthogonalization");
    if (!M) return 0;
    require(M->rows == A.n, "gmres: the preconditioner does not match the operator");
    require(M->shift == shift, "gmres: the preconditioner was built for another shift");
    return M->precision == MK_FP32 ? sizeof(float) : sizeof(double);
}

mk_solver_result run_fgmres(mk_generator_s &A, double shift, const precond *M, const double *b,
                            double *x, const mk_gmres_options &o, const krylov_space &K) {
    if (M && M->precision == MK_FP32) return fgmres<float>(A, shift, M, b, x, o, K);
    return fgmres<double>(A, shift, M, b, x, o, K);
}

// ---- the allocation-free workspace: one file-scope device array per GPU ----------------------
// Calls on one GPU share its array (the last call's contents stay in it): two calls that overlap
// on one GPU, from two host threads, race. Every call clears the small scalar arrays it uses.
constexpr std::size_t kScratchBytes = std::size_t{64} << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// the device address of the current GPU's scratch array, looked up once per device (with that
// device current) and kept; the host-side name `scratch` is only a handle for the lookup
unsigned char *scratch_address() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) throw device_error("scratch: device index out of range");
    static std::mutex mu;
    static void *addr[kScratchDevices] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw device_error("scratch: hipGetSymbolAddress returned no address");
        addr[dev] = p;
    }
    return static_cast<unsigned char *>(addr[dev]);
}

// carves 256-byte aligned arrays out of a base address (null base: sizes only)
struct carver {
    unsigned char *base = nullptr;
    std::size_t off = 0;
    template <typename T>
    T *take(std::size_t count) {
        off = (off + 255) / 256 * 256;
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
};

krylov_space carve_space(unsigned char *base, const mk_generator_s &A, i64 m, std::size_t zs,
                         std::size_t &bytes) {
    const i64 n = A.n;
    carver c{base};
    krylov_space K;
    K.ld = (n + 63) / 64 * 64;
    K.V = c.take<double>(static_cast<std::size_t>(K.ld) * (m + 1));
    K.Z = zs ? static_cast<void *>(c.take<unsigned char>(static_cast<std::size_t>(K.ld) * m * zs)) : nullptr;
    K.w = c.take<double>(n);
    K.partial = c.take<double>(static_cast<std::size_t>(gpu::reduction_blocks(n)) * (m + 2));
    K.hd = c.take<double>(3 * (m + 2));
    K.yd = c.take<double>(m + 1);
    K.rpart = c.take<double>(A.A.view.n_blocks + 1);
    K.scal = c.take<double>(2);
    bytes = c.off;
    return K;
}

mk_solver_result gmres_device_noalloc(mk_generator_s &A, double shift, const precond *M,
                                      const double *b, double *x, const mk_gmres_options &o) {
    const std::size_t zs = check_gmres(A, shift, M, o);
    const i64 m = gmres_restart(o, A.n);
    require(m <= gpu::kFusedBasis, "gmres (no allocation): restart must be at most 64");
    std::size_t bytes = 0;
    carve_space(nullptr, A, m, zs, bytes);
    if (bytes > kScratchBytes)
        throw std::invalid_argument("gmres (no allocation): the Krylov workspace needs " +
                                    std::to_string(bytes) + " bytes, above the " +
                                    std::to_string(kScratchBytes) + "-byte scratch array");
    krylov_space K = carve_space(scratch_address(), A, m, zs, bytes);
    // the scalar arrays start at zero on every call (the vectors are written before they are read)
    const hipStream_t s = stream();
    const i64 nb = gpu::reduction_blocks(A.n);
    gpu::fill<double>(nb * (m + 2), 0.0, K.partial, s);
    gpu::fill<double>(3 * (m + 2), 0.0, K.hd, s);
    gpu::fill<double>(m + 1, 0.0, K.yd, s);
    gpu::fill<double>(A.A.view.n_blocks + 1, 0.0, K.rpart, s);
    gpu::fill<double>(2, 0.0, K.scal, s);
    check_launch("scratch clear");
    double host[4 * (gpu::kFusedBasis + 2)] = {};
    K.host = host;
    K.hscal = nullptr;  // per-iteration scalars go by kernel arguments
    K.graphs = nullptr; // eager launches
    return run_fgmres(A, shift, M, b, x, o, K);
}
} // namespace

namespace mk::rt {
mk_solver_result gmres_device(mk_generator_s &A, double shift, const precond *M, const double *b,
                              double *x, const mk_gmres_options &o) {
    const std::size_t zs = check_gmres(A, shift, M, o);
    auto &ws = workspace();
    ws.ensure(A.n, gmres_restart(o, A.n), zs);
    if (ws.rpart_n < A.A.view.n_blocks + 1) {
        ws.rpart = dbuf();
        ws.rpart_n = A.A.view.n_blocks + 1;
        ws.rpart = dbuf(static_cast<std::size_t>(ws.rpart_n) * sizeof(double));
    }
    return run_fgmres(A, shift, M, b, x, o, cached_space(ws));
}
} // namespace mk::rt

namespace {
// re-invert the terminal blocks of every level for another shift (the level operators and
// transfers do not depend on it)
template <typename T>
void reshift_levels(std::vector<level_dev<T>> &L, double shift) {
    for (auto &lv : L) {
        if (lv.n_term == 0) continue;
        std::vector<T> inv(lv.h_tq.size());
        for (i64 b = 0; b < lv.n_term; ++b) {
            const i64 m = lv.h_tsize[b];
            std::vector<double> a(lv.h_tq.begin() + lv.h_toff[b], lv.h_tq.begin() + lv.h_toff[b] + m * m);
            for (i64 i = 0; i < m; ++i) a[i * m + i] += shift;
            std::vector<double> out;
            if (!mk::invert_dense(std::move(a), m, out))
                throw std::runtime_error("population_multigrid: singular block");
            std::copy(out.begin(), out.end(), inv.begin() + lv.h_toff[b]);
        }
        HIPCHK(hipMemcpyAsync(lv.tinv.p, inv.data(), inv.size() * sizeof(T), hipMemcpyHostToDevice,
                              stream()));
        drain();
    }
}
} // namespace

namespace {
// `sweeps` smoother sweeps on the finest level from the caller's x (population_multigrid::relax,
// multigrid.hpp:231-236), in the hierarchy's working precision; rhs and x in HBM, caller order
template <typename T>
void relax_top(mk_sector_mg_s &mg, std::vector<level_dev<T>> &L, const double *rhs64, double *x64,
               int sweeps, bool reverse) {
    auto &lv = L[0];
    require(!lv.direct && lv.A.n > 0, "relax: the level is solved directly, it has no smoother");
    const hipStream_t s = stream();
    const i64 n = lv.n;
    const double shift = mg.shift, omega = mg.opt.jacobi_weight;
    T *rhs = dp<T>(lv.rhs), *x = dp<T>(lv.x), *tmp = dp<T>(lv.xt), *res = dp<T>(lv.res);
    const u32 *perm = dp<u32>(lv.perm);
    if (perm) {
        gpu::permute_in<T, double>(n, perm, rhs64, rhs, s);
        gpu::permute_in<T, double>(n, perm, x64, x, s);
    } else {
        gpu::convert<double, T>(n, rhs64, rhs, s);
        gpu::convert<double, T>(n, x64, x, s);
    }
    check_launch("relax: input");
    if (lv.rb) {
        for (int k = 0; k < sweeps; ++k)
            for (int half = 0; half < 2; ++half) {
                const int c = reverse ? 1 - half : half;
                gpu::gs_half<T>(lv.Ac[c].view, shift, rhs, x, nullptr, s);
            }
    } else if (lv.ilu) {
        for (int k = 0; k < sweeps; ++k) {
            gpu::level_residual<T>(lv.A.view, shift, rhs, x, res, s);
            for (const auto &g : lv.fw_launch)
                gpu::ilu_forward<T>(lv.A.view, dp<u32>(lv.fw_rows), dp<u32>(lv.fw_ptr), g.g0, g.g1,
                                    g.wide_rows, dp<T>(lv.pinv), res, tmp, s);
            for (const auto &g : lv.bw_launch)
                gpu::ilu_backward<T>(lv.A.view, dp<u32>(lv.bw_rows), dp<u32>(lv.bw_ptr), g.g0, g.g1,
                                     g.wide_rows, dp<T>(lv.pinv), tmp, x, s);
        }
    } else if (lv.n_agg > 0) {
        block_smooth<T>(mg, lv, 0, rhs, x, tmp, sweeps, reverse, false);
    } else {
        for (int k = 0; k < sweeps; ++k) {
            gpu::jacobi_sweep<T>(lv.A.view, shift, omega, rhs, x, tmp, s);
            std::swap(x, tmp);
        }
    }
    check_launch("relax: sweeps");
    const T *nat = x;
    if (perm) {
        T *other = x == res ? tmp : res;
        gpu::permute_out<T>(n, perm, x, other, s);
        nat = other;
    }
    gpu::convert<T, double>(n, nat, x64, s);
    check_launch("relax: output");
}

// the stored entries of a generator in gather form, decoded on the host from its device arrays:
// row i lists (column j, rate of j -> i), columns ascending
void download_gather(const mk_generator_s &G, std::vector<int64_t> &ptr, std::vector<int64_t> &col,
                     std::vector<double> &val) {
    const auto &A = G.A;
    const i64 n = A.n;
    ptr.assign(n + 1, 0);
    col.clear();
    val.clear();
    auto fetch = [](const dbuf &b, auto &v) {
        v.resize(b.bytes / sizeof(v[0]));
        if (b.bytes) HIPCHK(hipMemcpy(v.data(), b.p, v.size() * sizeof(v[0]), hipMemcpyDeviceToHost));
    };
    if (A.coded) {
        std::vector<u32> soff, ent;
        std::vector<std::uint8_t> rlen;
        std::vector<double> vdict;
        fetch(A.soff, soff);
        fetch(A.ent, ent);
        fetch(A.rlen, rlen);
        fetch(A.vdict, vdict);
        for (i64 i = 0; i < n; ++i) {
            const u64 base = soff[i >> 6] + static_cast<u64>(i & 63);
            for (u32 k = 0; k < rlen[i]; ++k) {
                const u32 e = ent[base + static_cast<u64>(k) * gpu::kSlice];
                col.push_back(e & gpu::kColMask);
                val.push_back(vdict[e >> gpu::kColBits]);
            }
            ptr[i + 1] = static_cast<int64_t>(col.size());
        }
    } else {
        std::vector<u32> p, c;
        fetch(A.ptr, p);
        fetch(A.col, c);
        fetch(A.val, val);
        for (i64 i = 0; i <= n; ++i) ptr[i] = p[i];
        col.assign(c.begin(), c.end());
    }
}
} // namespace

namespace mk::rt {
void finish_generator(mk_generator_s &G, mk::generator_csr &&g) {
    G.n = g.n;
    G.n_jumps = g.n_jumps;
    // ||Q||_1 = max_j (exit_j + sum of j's in-space jump rates)
    std::vector<double> out(g.exit.begin(), g.exit.end());
    for (i64 i = 0; i < g.n; ++i)
        for (u64 p = g.ptr[i]; p < g.ptr[i + 1]; ++p) out[g.col[p]] += g.val[p];
    G.norm1 = out.empty() ? 0.0 : *std::max_element(out.begin(), out.end());
    G.A = upload_csr<double>(g);
    G.exit = std::move(g.exit);
    G.leak = std::move(g.leak);
}
} // namespace mk::rt

// ====================================================================== C ABI
extern "C" {

const char *mk_last_error(void) { return last_error().c_str(); }

mk_status mk_graphs_enable(int32_t on) {
    graphs_switch().store(on ? 1 : 0);
    return MK_OK;
}

mk_status mk_set_xcd_chunk(int32_t chunk) {
    return guarded([&] {
        require(chunk >= 0 && chunk <= 4096, "xcd chunk: 0 to 4096 row blocks");
        gpu::set_xcd_chunk(chunk);
    });
}

mk_status mk_get_xcd_chunk(int32_t *chunk) {
    return guarded([&] {
        require(chunk != nullptr, "xcd chunk: bad argument");
        *chunk = gpu::xcd_chunk();
    });
}

mk_status mk_device_setup_enable(int32_t on) {
    device_setup_switch().store(on ? 1 : 0);
    return MK_OK;
}

mk_status mk_fused_first_enable(int32_t on) {
    fused_first_switch().store(on ? 1 : 0);
    return MK_OK;
}

mk_status mk_coarse_fuse_enable(int32_t on) {
    coarse_fuse_switch().store(on ? 1 : 0);
    return MK_OK;
}

mk_status mk_tiny_gmres_enable(int32_t on) {
    tiny_gmres_switch().store(on ? 1 : 0);
    return MK_OK;
}

mk_status mk_tail_lds_enable(int32_t on) {
    tail_lds_switch().store(on ? 1 : 0);
    return MK_OK;
}

mk_status mk_fused_tail_enable(int32_t on) {
    tail_switch().store(on ? 1 : 0);
    return MK_OK;
}

void mk_default_multigrid_options(mk_multigrid_options *o) {
    o->pre_sweeps = 2;
    o->post_sweeps = 2;
    o->jacobi_weight = 2.0 / 3.0;
    o->coarsest_size = 4;
    o->max_levels = 32;
    o->coarse_cycles = 1;
    o->sweep_growth = 0;
    o->smoother = MK_JACOBI;
    o->precision = MK_FP64;
    o->cycle_levels = 0;
}

mk_status mk_box_create(int64_t n_voxels, int32_t n_species, const int32_t *max_totals,
                        mk_box *out) {
    return guarded([&] {
        require(out && max_totals && n_species > 0, "population_box: bad arguments");
        auto b = std::make_unique<mk_box_s>();
        b->space = mk::population_space::box(n_voxels, std::vector<int>(max_totals, max_totals + n_species));
        *out = b.release();
    });
}

mk_status mk_box_destroy(mk_box box) {
    delete box;
    return MK_OK;
}

mk_status mk_box_size(mk_box box, int64_t *size) {
    return guarded([&] { *size = box->space.size(); });
}

mk_status mk_box_n_sectors(mk_box box, int64_t *n) {
    return guarded([&] { *n = box->space.n_sectors(); });
}

mk_status mk_box_sector(mk_box box, int64_t k, int64_t *begin, int64_t *end, int32_t *totals) {
    return guarded([&] {
        require(k >= 0 && k < box->space.n_sectors(), "population_box: sector out of range");
        if (begin) *begin = box->space.begin(k);
        if (end) *end = box->space.end(k);
        if (totals)
            for (int s = 0; s < box->space.species(); ++s) totals[s] = box->space.totals(k)[s];
    });
}

mk_status mk_box_state(mk_box box, int64_t i, int32_t *state) {
    return guarded([&] {
        require(i >= 0 && i < box->space.size(), "population_box: index out of range");
        box->space.state(i, state);
    });
}

mk_status mk_box_find(mk_box box, const int32_t *state, int64_t *index) {
    return guarded([&] { *index = box->space.find(state); });
}

mk_status mk_sector_space_create(int64_t n_voxels, int32_t n_species, const int32_t *totals,
                                mk_box *out) {
    return guarded([&] {
        require(out && totals && n_species > 0, "population_sector: bad arguments");
        auto b = std::make_unique<mk_box_s>();
        b->space = mk::population_space::sector(n_voxels, std::vector<int>(totals, totals + n_species));
        *out = b.release();
    });
}

mk_status mk_domain_create_grid(int32_t dim, const int64_t *dims, const double *spacing,
                                const double *volumes, int32_t n_species,
                                const double *diffusivities, mk_domain *out) {
    return guarded([&] {
        require(dim >= 1 && dim <= 3 && dims && spacing && n_species > 0 && diffusivities,
                "voxel_domain: bad grid arguments");
        auto d = std::make_unique<mk_domain_s>();
        d->dom = mk::voxel_domain::grid(std::vector<i64>(dims, dims + dim),
                                        std::vector<double>(spacing, spacing + dim), volumes,
                                        std::vector<double>(diffusivities, diffusivities + n_species));
        *out = d.release();
    });
}

mk_status mk_domain_create_graph(int64_t n_voxels, const int64_t *offsets, const int64_t *to,
                                 const double *weight, const double *volumes, int32_t n_species,
                                 const double *diffusivities, mk_domain *out) {
    return guarded([&] {
        require(n_species > 0 && diffusivities, "voxel_domain: bad graph arguments");
        auto d = std::make_unique<mk_domain_s>();
        d->dom = mk::voxel_domain::graph(n_voxels, offsets, to, weight, volumes,
                                         std::vector<double>(diffusivities, diffusivities + n_species));
        *out = d.release();
    });
}

mk_status mk_domain_destroy(mk_domain d) {
    delete d;
    return MK_OK;
}



mk_status mk_generator_create_rdme(mk_box box, mk_domain domain, int32_t n_reactions,
                                   const int32_t *changes, const double *rates, int32_t boundary,
                                   mk_generator *out) {
    return guarded([&] {
        require(box && domain && out, "generator: bad arguments");
        require(n_reactions >= 0 && (n_reactions == 0 || (changes && rates)),
                "generator: reactions need changes and rate tables");
        const auto &space = box->space;
        const int S = space.species();
        require(domain->dom.species() == S && domain->dom.voxels == space.voxels(),
                "generator: the domain does not match the box");
        mk::reaction_tables rt;
        rt.n = n_reactions;
        rt.stride.assign(S, 1);
        for (int s = S; s-- > 0;) {
            rt.stride[s] = rt.T;
            rt.T *= space.caps()[s] + 1;
        }
        if (n_reactions > 0) {
            rt.changes.assign(changes, changes + static_cast<i64>(n_reactions) * S);
            rt.rates.assign(rates, rates + static_cast<i64>(n_reactions) * space.voxels() * rt.T);
        }
        auto G = std::make_unique<mk_generator_s>();
        const bool absorbing = boundary == MK_ABSORBING;
        if (!(device_setup_switch().load() != 0 &&
              assemble_generator_device(space, domain->dom, rt, absorbing, *G)))
            finish_generator(*G, mk::assemble_generator(space, domain->dom, rt, absorbing));
        *out = G.release();
    });
}

mk_status mk_generator_create_jumps(int64_t n, const int64_t *first, const int64_t *to,
                                    const double *rate, const double *exit, const double *leak,
                                    mk_generator *out) {
    return guarded([&] {
        require(first && exit && leak && out, "generator: bad arguments");
        auto G = std::make_unique<mk_generator_s>();
        finish_generator(*G, mk::generator_from_jumps(n, first, to, rate, exit, leak));
        *out = G.release();
    });
}

mk_status mk_generator_destroy(mk_generator g) {
    delete g;
    return MK_OK;
}

mk_status mk_generator_size(mk_generator g, int64_t *rows, int64_t *n_jumps) {
    return guarded([&] {
        if (rows) *rows = g->n;
        if (n_jumps) *n_jumps = g->n_jumps;
    });
}

mk_status mk_generator_entries(mk_generator g, int64_t *entries, int64_t *ptr, int64_t *col,
                               double *rate) {
    return guarded([&] {
        require(g && entries, "generator: bad arguments");
        std::vector<int64_t> p, c;
        std::vector<double> v;
        download_gather(*g, p, c, v);
        *entries = static_cast<int64_t>(c.size());
        if (ptr) std::copy(p.begin(), p.end(), ptr);
        if (col) std::copy(c.begin(), c.end(), col);
        if (rate) std::copy(v.begin(), v.end(), rate);
    });
}

mk_status mk_generator_rates(mk_generator g, double *exit, double *leak) {
    return guarded([&] {
        if (exit) std::copy(g->exit.begin(), g->exit.end(), exit);
        if (leak) std::copy(g->leak.begin(), g->leak.end(), leak);
    });
}

namespace {
void stage(mk_generator_s &g) {
    if (!g.d_x.p) g.d_x = dbuf(std::max<i64>(g.n, 1) * sizeof(double));
    if (!g.d_y.p) g.d_y = dbuf(std::max<i64>(g.n, 1) * sizeof(double));
}
} // namespace

mk_status mk_shifted_apply_device(mk_generator g, double shift, const double *d_x, double *d_y) {
    return guarded([&] {
        prof().run(K_SPMV, op_bytes(g->A, 8, 8, 1), [&] {
            gpu::shifted_apply<double>(g->A.view, shift, d_x, nullptr, d_y, nullptr, stream());
        });
    });
}

mk_status mk_shifted_apply(mk_generator g, double shift, const double *x, double *y) {
    return guarded([&] {
        stage(*g);
        HIPCHK(hipMemcpy(g->d_x.p, x, g->n * sizeof(double), hipMemcpyHostToDevice));
        prof().run(K_SPMV, op_bytes(g->A, 8, 8, 1), [&] {
            gpu::shifted_apply<double>(g->A.view, shift,
                                       dp<double>(g->d_x), nullptr, dp<double>(g->d_y), nullptr,
                                       stream());
        });
        drain();
        HIPCHK(hipMemcpy(y, g->d_y.p, g->n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

mk_status mk_generator_apply(mk_generator g, const double *x, double *y) {
    // Q x = -(0 I - Q) x
    const mk_status st = mk_shifted_apply(g, 0.0, x, y);
    if (st == MK_OK)
        for (i64 i = 0; i < g->n; ++i) y[i] = -y[i];
    return st;
}

mk_status mk_sector_mg_create(mk_domain domain, mk_box box, double shift,
                              const mk_multigrid_options *options, mk_sector_mg *out) {
    return guarded([&] {
        require(domain && box && out, "sector_multigrid: bad arguments");
        mk_multigrid_options o;
        mk_default_multigrid_options(&o);
        if (options) o = *options;
        require(o.smoother == MK_JACOBI || o.smoother == MK_AGGREGATE_BLOCK || o.smoother == MK_RED_BLACK ||
                    o.smoother == MK_ILU0,
                "sector_multigrid: unknown smoother");
        require(o.precision == MK_FP64 || o.precision == MK_FP32, "sector_multigrid: unknown precision");
        require(o.pre_sweeps >= 0 && o.post_sweeps >= 0 && o.coarse_cycles >= 1 && o.cycle_levels >= 0,
                "sector_multigrid: bad sweep or cycle counts");
        if (!(shift > 0.0)) throw std::invalid_argument("population_multigrid: shift must be positive");
        const auto &space = box->space;
        require(domain->dom.species() == space.species() && domain->dom.voxels == space.voxels(),
                "sector_multigrid: the domain does not match the box");
        auto mg = std::make_unique<mk_sector_mg_s>();
        mg->opt.pre_sweeps = o.pre_sweeps;
        mg->opt.post_sweeps = o.post_sweeps;
        mg->opt.jacobi_weight = o.jacobi_weight;
        mg->opt.coarsest_size = o.coarsest_size;
        mg->opt.max_levels = o.max_levels;
        mg->opt.coarse_cycles = o.coarse_cycles;
        mg->opt.cycle_levels = o.cycle_levels;
        mg->opt.sweep_growth = o.sweep_growth;
        mg->opt.smoother = o.smoother;
        mg->shift = shift;
        mg->precision = o.precision;
        mg->rows = space.size();
        mg->n_sectors = space.n_sectors();
        // level operators and transfers of at least 4096 states are built on the device (the
        // same arrays)
        mk::level_hooks on_device;
        if (device_setup_switch().load() != 0) {
            on_device.assemble = [](const mk::population_space &sp, const mk::voxel_domain &d) {
                const mk::reaction_tables none;
                mk::generator_csr g;
                if (sp.size() < 4096 || !assemble_csr_device(sp, d, none, true, g))
                    g = mk::assemble_generator(sp, d, none, true);
                return g;
            };
            on_device.transfer = transfer_device;
        }
        std::vector<mk::sector_hierarchy> H;
        for (i64 k = 0; k < space.n_sectors(); ++k)
            H.push_back(mk::build_sector_hierarchy(domain->dom, space.totals(k), shift, mg->opt, on_device));
        if (o.precision == MK_FP32) {
            build_levels<float>(*mg, H, mg->L32);
        } else {
            build_levels<double>(*mg, H, mg->L64);
        }
        if (!space.is_box() && space.n_sectors() == 1)
            mg->host_levels = std::make_shared<mk::sector_hierarchy>(std::move(H[0]));
        *out = mg.release();
    });
}

mk_status mk_sector_mg_set_shift(mk_sector_mg mg, double shift) {
    return guarded([&] {
        require(mg != nullptr, "sector_multigrid: bad arguments");
        if (!(shift > 0.0)) throw std::invalid_argument("population_multigrid: shift must be positive");
        require(mg->opt.smoother == MK_JACOBI || mg->opt.smoother == MK_RED_BLACK,
                "sector_multigrid: the aggregate-block and ILU(0) smoothers' factors depend on the shift; "
                "build a new one");
        if (mg->precision == MK_FP32) {
            reshift_levels<float>(mg->L32, shift);
        } else {
            reshift_levels<double>(mg->L64, shift);
        }
        refresh_tail_image(*mg);
        mg->shift = shift;
    });
}

mk_status mk_sector_mg_level(mk_sector_mg mg, int32_t level, int64_t *rows, int64_t *jumps,
                             int32_t *terminal) {
    return guarded([&] {
        require(mg && mg->host_levels, "population_multigrid: the hierarchy is not over one population_sector");
        const auto &H = mg->host_levels->levels;
        require(level >= 0 && level < static_cast<int32_t>(H.size()), "population_multigrid: no such level");
        const auto &lv = H[level];
        if (rows) *rows = lv.rows;
        if (jumps) *jumps = lv.A.n_jumps;
        if (terminal) *terminal = lv.terminal ? 1 : 0;
    });
}

mk_status mk_sector_mg_transfer(mk_sector_mg mg, int32_t level, int64_t *coarse_of, double *weight) {
    return guarded([&] {
        require(mg && mg->host_levels && coarse_of && weight,
                "population_multigrid: the hierarchy is not over one population_sector");
        const auto &H = mg->host_levels->levels;
        require(level >= 0 && level + 1 < static_cast<int32_t>(H.size()) && !H[level].terminal,
                "population_multigrid: the level has no coarse level");
        std::copy(H[level].coarse_of.begin(), H[level].coarse_of.end(), coarse_of);
        std::copy(H[level].weight.begin(), H[level].weight.end(), weight);
    });
}

mk_status mk_sector_mg_relax_device(mk_sector_mg mg, const double *d_rhs, double *d_x,
                                    int32_t sweeps, int32_t reverse) {
    return guarded([&] {
        require(mg && d_rhs && d_x && sweeps >= 0, "population_multigrid: bad arguments");
        if (mg->precision == MK_FP32) {
            relax_top<float>(*mg, mg->L32, d_rhs, d_x, sweeps, reverse != 0);
        } else {
            relax_top<double>(*mg, mg->L64, d_rhs, d_x, sweeps, reverse != 0);
        }
        drain();
    });
}

mk_status mk_sector_mg_relax(mk_sector_mg mg, const double *rhs, double *x, int32_t sweeps,
                             int32_t reverse) {
    return guarded([&] {
        require(mg && rhs && x && sweeps >= 0, "population_multigrid: bad arguments");
        const std::size_t bytes = static_cast<std::size_t>(mg->rows) * sizeof(double);
        if (!mg->d_b.p) mg->d_b = dbuf(std::max<std::size_t>(bytes, 8));
        if (!mg->d_x64.p) mg->d_x64 = dbuf(std::max<std::size_t>(bytes, 8));
        HIPCHK(hipMemcpy(mg->d_b.p, rhs, bytes, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(mg->d_x64.p, x, bytes, hipMemcpyHostToDevice));
        const mk_status st = mk_sector_mg_relax_device(mg, dp<double>(mg->d_b), dp<double>(mg->d_x64),
                                                       sweeps, reverse);
        if (st != MK_OK) throw std::runtime_error(mk_last_error());
        HIPCHK(hipMemcpy(x, mg->d_x64.p, bytes, hipMemcpyDeviceToHost));
    });
}

mk_status mk_sector_mg_destroy(mk_sector_mg mg) {
    delete mg;
    return MK_OK;
}

mk_status mk_sector_mg_info(mk_sector_mg mg, int64_t *rows, int64_t *n_sectors,
                            int64_t *max_levels) {
    return guarded([&] {
        if (rows) *rows = mg->rows;
        if (n_sectors) *n_sectors = mg->n_sectors;
        if (max_levels) *max_levels = mg->max_levels;
    });
}

mk_status mk_sector_mg_apply_device(mk_sector_mg mg, const double *d_b, double *d_x) {
    return guarded([&] {
        if (mg->precision == MK_FP32) {
            if (!mg->d_out.p) mg->d_out = dbuf(std::max<i64>(mg->rows, 1) * sizeof(float));
            mg_apply_device(*mg, d_b, mg->d_out.p);
            prof().run(K_OTHER, mg->rows * 12.0, [&] {
                gpu::convert<float, double>(mg->rows, dp<float>(mg->d_out), d_x, stream());
            });
        } else {
            mg_apply_device(*mg, d_b, d_x);
        }
    });
}

mk_status mk_sector_mg_apply(mk_sector_mg mg, const double *b, double *x) {
    return guarded([&] {
        const std::size_t bytes = std::max<i64>(mg->rows, 1) * sizeof(double);
        if (!mg->d_b.p) mg->d_b = dbuf(bytes);
        if (!mg->d_x64.p) mg->d_x64 = dbuf(bytes);
        HIPCHK(hipMemcpy(mg->d_b.p, b, mg->rows * sizeof(double), hipMemcpyHostToDevice));
        if (mg->precision == MK_FP32) {
            if (!mg->d_out.p) mg->d_out = dbuf(std::max<i64>(mg->rows, 1) * sizeof(float));
            mg_apply_device(*mg, dp<double>(mg->d_b), mg->d_out.p);
            prof().run(K_OTHER, mg->rows * 12.0, [&] {
                gpu::convert<float, double>(mg->rows, dp<float>(mg->d_out), dp<double>(mg->d_x64), stream());
            });
        } else {
            mg_apply_device(*mg, dp<double>(mg->d_b), dp<double>(mg->d_x64));
        }
        drain();
        HIPCHK(hipMemcpy(x, mg->d_x64.p, mg->rows * sizeof(double), hipMemcpyDeviceToHost));
    });
}

mk_status mk_gmres_device(mk_generator gen, double shift, mk_sector_mg mg, const double *d_b,
                          double *d_x, const mk_gmres_options *options, mk_solver_result *result) {
    return guarded([&] {
        require(gen && options && result, "gmres: bad arguments");
        require(options->max_iterations >= 0 && options->restart >= 1, "gmres: bad options");
        *result = gmres_device(*gen, shift, mg, d_b, d_x, *options);
        drain();
    });
}

mk_status mk_gmres_device_noalloc(mk_generator gen, double shift, mk_sector_mg mg,
                                  const double *d_b, double *d_x, const mk_gmres_options *options,
                                  mk_solver_result *result) {
    return guarded([&] {
        require(gen && d_b && d_x && options && result, "gmres: bad arguments");
        require(options->max_iterations >= 0 && options->restart >= 1, "gmres: bad options");
        if (mg) {
            const precond M = sector_precond(mg);
            *result = gmres_device_noalloc(*gen, shift, &M, d_b, d_x, *options);
        } else {
            *result = gmres_device_noalloc(*gen, shift, nullptr, d_b, d_x, *options);
        }
        drain();
    });
}

mk_status mk_gmres(mk_generator gen, double shift, mk_sector_mg mg, const double *b, double *x,
                   const mk_gmres_options *options, mk_solver_result *result) {
    return guarded([&] {
        require(gen && options && result, "gmres: bad arguments");
        require(options->max_iterations >= 0 && options->restart >= 1, "gmres: bad options");
        stage(*gen);
        HIPCHK(hipMemcpy(gen->d_x.p, x, gen->n * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(gen->d_y.p, b, gen->n * sizeof(double), hipMemcpyHostToDevice));
        *result = gmres_device(*gen, shift, mg, dp<double>(gen->d_y), dp<double>(gen->d_x), *options);
        drain();
        HIPCHK(hipMemcpy(x, gen->d_x.p, gen->n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

mk_status mk_device_alloc(int64_t bytes, void **d_ptr) {
    return guarded([&] {
        require(bytes >= 0 && d_ptr, "device_alloc: bad arguments");
        *d_ptr = nullptr;
        if (bytes > 0) HIPCHK(hipMalloc(d_ptr, bytes));
    });
}

mk_status mk_device_free(void *d_ptr) {
    return guarded([&] {
        if (d_ptr) HIPCHK(hipFree(d_ptr));
    });
}

mk_status mk_copy_to_device(void *d_dst, const void *src, int64_t bytes) {
    return guarded([&] { HIPCHK(hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice)); });
}

mk_status mk_copy_to_host(void *dst, const void *d_src, int64_t bytes) {
    return guarded([&] {
        drain();
        HIPCHK(hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost));
    });
}

mk_status mk_device_zero(void *d_ptr, int64_t bytes) {
    return guarded([&] {
        require(bytes >= 0, "device_zero: negative size");
        if (bytes > 0) HIPCHK(hipMemsetAsync(d_ptr, 0, bytes, stream()));
    });
}

mk_status mk_device_copy(void *d_dst, const void *d_src, int64_t bytes) {
    return guarded([&] {
        require(bytes >= 0 && bytes % 16 == 0, "device_copy: size must be a multiple of 16 bytes");
        require((reinterpret_cast<std::uintptr_t>(d_dst) | reinterpret_cast<std::uintptr_t>(d_src)) % 16 == 0,
                "device_copy: pointers must be 16-byte aligned");
        prof().run(K_OTHER, 2.0 * bytes, [&] { gpu::copy16(bytes / 16, d_src, d_dst, stream()); });
    });
}

mk_status mk_synchronize(void) {
    return guarded([&] { drain(); });
}

mk_status mk_profile_enable(int32_t mask) {
    return guarded([&] {
        drain();
        prof().mask = static_cast<std::uint32_t>(mask) & ((1u << K_COUNT) - 1u);
    });
}

mk_status mk_profile_reset(void) {
    return guarded([&] {
        drain();
        prof().reset();
    });
}

mk_status mk_profile_read(const char *kernel, int64_t *launches, double *milliseconds,
                          double *bytes) {
    return guarded([&] {
        drain();
        for (int c = 0; c < K_COUNT; ++c) {
            if (std::strcmp(kernel, kclass_names[c]) == 0) {
                if (launches) *launches = prof().launches[c];
                if (milliseconds) *milliseconds = prof().ms[c];
                if (bytes) *bytes = prof().bytes[c];
                return;
            }
        }
        throw std::invalid_argument("profile: unknown kernel class");
    });
}

} // extern "C"
