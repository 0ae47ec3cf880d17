// Circuit templates of the lane engine (host): node numbering, the frozen
// MNA pattern, slot tapes, the shared column order. See engine.h Template.
#include <algorithm>
#include <cstring>

#include "engine.h"
#include "wave.h"

namespace esp {

uint32_t Template::node(const std::string &name)
{
    if (name == "0" || name == "gnd") return 0;
    auto it = names.find(name);
    if (it != names.end()) return it->second;
    uint32_t r = n++;
    cur_flag.push_back(0);
    names.emplace(name, r);
    return r;
}

int Template::add(int kind, const uint32_t *ports, const double *params, uint32_t *inst)
{
    if (frozen || kind < 0 || kind >= K_NKIND) return -1;
    for (int p = 0; p < KIND_NPORT[kind]; p++)
        if (ports[p] >= n) return -1;
    bool seen = false;
    for (int i = 0; i < n_order; i++) seen = seen || order[i] == kind;
    if (!seen) order[n_order++] = kind;
    if (kind_is_va(kind)) {
        // a compiled kind's collapse, decided by its setup on these
        // parameters (the generated collapse pairs)
        int64_t rows[32];
        if (va_collapse_rows(kind, params, ports, rows) != 0) return -1;
        const size_t base = nodes[kind].size();
        for (int u = 0; u < KIND_NU[kind]; u++) {
            if (rows[u] >= 0) {
                nodes[kind].push_back((uint32_t)rows[u]);
            } else if (rows[u] == -1) {
                nodes[kind].push_back(n++);
                cur_flag.push_back(0);
            } else {
                nodes[kind].push_back(nodes[kind][base + (size_t)(-2 - rows[u])]);
            }
        }
        for (int k = 0; k < KIND_UK[kind]; k++) uprm[kind].push_back(params[k]);
        *inst = count[kind]++;
        return 0;
    }
    for (int p = 0; p < KIND_NPORT[kind]; p++) nodes[kind].push_back(ports[p]);
    // Internal unknowns take fresh rows in unknown order unless the kind's
    // collapse maps them onto a port (frontend/builder.zig:258-263): a
    // diode's ai is its anode when RS is not positive (diode.va:364-371).
    for (int u = KIND_NPORT[kind]; u < KIND_NU[kind]; u++) {
        if (kind == K_DIODE && !(params[DIO_RS] > 0.0)) {
            nodes[kind].push_back(ports[0]);
            continue;
        }
        nodes[kind].push_back(n++);
        cur_flag.push_back(kind == K_VSRC ? 1 : 0);
    }
    for (int k = 0; k < KIND_UK[kind]; k++) uprm[kind].push_back(params[k]);
    *inst = count[kind]++;
    return 0;
}

uint32_t Template::find_slot(uint32_t row, uint32_t col) const
{
    uint32_t lo = col_ptr[col], hi = col_ptr[col + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (row_idx[mid] < row) lo = mid + 1;
        else hi = mid;
    }
    return (lo < col_ptr[col + 1] && row_idx[lo] == row) ? lo : ESP_NONE;
}

bool is_tridiag(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx)
{
    if (n < 3) return false;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t r = row_idx[p];
            if ((r >= j ? r - j : j - r) > 1) return false;
        }
    return true;
}

int Template::freeze()
{
    if (frozen) return 0;
    // PatternBuilder (device/abi.zig:573-643): (col << 32 | row) keys, every
    // diagonal first, sorted and deduplicated per column.
    std::vector<uint64_t> keys;
    keys.reserve(n);
    for (uint32_t i = 0; i < n; i++) keys.push_back(((uint64_t)i << 32) | i);
    for (int t = 0; t < n_order; t++) {
        const int kind = order[t], nu = KIND_NU[kind];
        for (uint32_t id = 0; id < count[kind]; id++) {
            const uint32_t *nd = &nodes[kind][(size_t)id * nu];
            for (int ru = 0; ru < nu; ru++)
                for (int cu = 0; cu < nu; cu++)
                    if ((kind_pmask(kind, ru) >> cu & 1) && nd[ru] != 0 && nd[cu] != 0)
                        keys.push_back(((uint64_t)nd[cu] << 32) | nd[ru]);
        }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    col_ptr.assign(n + 1, 0);
    row_idx.resize(keys.size());
    for (size_t i = 0; i < keys.size(); i++) {
        row_idx[i] = (uint32_t)keys[i];
        col_ptr[(keys[i] >> 32) + 1]++;
    }
    for (uint32_t j = 0; j < n; j++) col_ptr[j + 1] += col_ptr[j];
    nnz = (uint32_t)keys.size();
    diag.resize(n);
    for (uint32_t i = 0; i < n; i++) diag[i] = find_slot(i, i);
    cur_row.assign(cur_flag.begin(), cur_flag.begin() + n);
    for (int kind = 0; kind < K_NKIND; kind++) {
        uint32_t ks = 0;
        for (int ru = 0; ru < KIND_NU[kind]; ru++) ks += (uint32_t)__builtin_popcount(kind_pmask(kind, ru));
        kslots[kind] = ks;
        slots[kind].clear();
        const int nu = KIND_NU[kind];
        for (uint32_t id = 0; id < count[kind]; id++) {
            const uint32_t *nd = &nodes[kind][(size_t)id * nu];
            for (int ru = 0; ru < nu; ru++)
                for (int cu = 0; cu < nu; cu++)
                    if (kind_pmask(kind, ru) >> cu & 1)
                        slots[kind].push_back((nd[ru] == 0 || nd[cu] == 0) ? nnz : find_slot(nd[ru], nd[cu]));
        }
    }
    q.resize(n);
    if (column_order(n, col_ptr.data(), row_idx.data(), q.data()) != 0) return -1;
    // direct.zig:233-234: at least nnz(A).
    uint64_t df = 2 * (chol_count(n, col_ptr.data(), row_idx.data(), q.data(), UINT64_MAX) - n);
    diag_fill = std::max<uint64_t>(df, nnz);
    tridiagonal = is_tridiag(n, col_ptr.data(), row_idx.data());
    tri_pos.assign(3 * (size_t)n, ESP_NONE);
    if (tridiagonal)
        for (uint32_t j = 0; j < n; j++)
            for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
                uint32_t r = row_idx[p];
                if (r == j) tri_pos[n + j] = p;
                else if (r == j + 1) tri_pos[j + 1] = p;
                else if (j > 0 && r + 1 == j) tri_pos[2 * n + r] = p;
            }
    frozen = true;
    return 0;
}

// ngspice's placement of the stepping rungs' diagonal gmin
// (src/analysis/dc/op.zig:711-790 gminStamps, after spMNA_Preorder of
// sputils.c and LoadGmin of spsmp.c): every column with no diagonal
// element (a voltage source's branch, whose diagonal nothing stamps) is
// swapped with a column holding a symmetric pair of +-1 entries, lone
// pairs first, then the first column with several; each row's gmin then
// lands on the entry its column's position holds. A grounded source's
// gmin so goes onto its two incidence entries rather than its node's
// diagonal. The reference reads |value| == 1 off an assembled iterate; in
// the lane kinds a branch column and its row hold nothing but the
// source's own +-1 incidence (no other kind stamps a branch), so the test
// is structural here: `unit` marks those entries.
void Template::gmin_stamps(std::vector<uint32_t> &out) const
{
    std::vector<uint8_t> unit(nnz, 0);
    for (uint32_t id = 0; id < count[K_VSRC]; id++) {
        const uint32_t *nd = &nodes[K_VSRC][(size_t)id * 3];
        if (nd[0] == nd[1]) continue;  // +1 and -1 on one entry sum to 0
        for (int t = 0; t < 2; t++) {
            if (nd[t] == 0) continue;
            unit[find_slot(nd[t], nd[2])] = 1;
            unit[find_slot(nd[2], nd[t])] = 1;
        }
    }
    std::vector<uint32_t> perm(n), dcol(n);
    for (uint32_t i = 0; i < n; i++) {
        perm[i] = i;
        dcol[i] = cur_row[i] ? ESP_NONE : i;
    }
    // the first +-1 entry of position j's column whose transpose is +-1 too,
    // and how many such (counting stops at 2)
    auto twins = [&](uint32_t j, uint32_t &row) {
        uint32_t cnt = 0;
        const uint32_t col = perm[j];
        for (uint32_t p = col_ptr[col]; p < col_ptr[col + 1]; p++) {
            const uint32_t r = row_idx[p];
            if (r == 0 || !unit[p]) continue;
            const uint32_t q = find_slot(j, perm[r]);
            if (q == ESP_NONE || !unit[q]) continue;
            if (++cnt == 1) row = r;
            else break;
        }
        return cnt;
    };
    auto swap = [&](uint32_t j, uint32_t r) {
        dcol[j] = perm[r];
        dcol[r] = perm[j];
        std::swap(perm[j], perm[r]);
    };
    uint32_t start = 1;
    for (;;) {
        bool swapped = false, again = false;
        for (uint32_t j = start; j < n; j++) {
            if (dcol[j] != ESP_NONE) continue;
            uint32_t r = 0;
            const uint32_t c = twins(j, r);
            if (c == 1) {
                swap(j, r);
                swapped = true;
            } else if (c > 1 && !again) {
                again = true;
                start = j;
            }
        }
        if (!again) break;
        for (uint32_t j = start; !swapped && j < n; j++) {
            if (dcol[j] != ESP_NONE) continue;
            uint32_t r = 0;
            if (twins(j, r) > 0) {
                swap(j, r);
                swapped = true;
            }
        }
        if (!swapped) break;
    }
    out.assign(2 * (size_t)n, ESP_NONE);
    for (uint32_t i = 1; i < n; i++) {
        if (dcol[i] == ESP_NONE) continue;
        out[i] = dcol[i];
        out[n + i] = find_slot(i, dcol[i]);
    }
}

}  // namespace esp


namespace esp {

void build_wave(const Template &T, uint32_t cap, WaveHost &H)
{
    std::vector<InstEntry> inst;
    uint32_t stage = 0, prm = 0, lim = 0, tape = 0;
    // prepared constants: the hand-written kinds' first (copied into each
    // workgroup's LDS), then the compiled kinds' (read from HBM: hundreds of
    // constants per instance that would crowd the lane block out of LDS)
    uint32_t prm_lds = 0;
    for (int t = 0; t < T.n_order; t++)
        if (!kind_is_va(T.order[t])) prm_lds += T.count[T.order[t]] * (uint32_t)KIND_PK[T.order[t]];
    uint32_t prm_va = prm_lds;
    std::vector<std::vector<uint32_t>> rows(T.n), slots(T.nnz), qrows(T.n), cslots(T.nnz);
    H.nodes_flat.clear();
    H.slots_flat.clear();
    H.node_off.clear();
    H.slot_off.clear();
    for (int t = 0; t < T.n_order; t++) {
        const int kind = T.order[t], nu = KIND_NU[kind];
        for (uint32_t id = 0; id < T.count[kind]; id++) {
            InstEntry e{};
            e.kind = (uint8_t)kind;
            e.id = id;
            e.stage_res = stage;
            stage += nu;
            e.stage_jac = stage;
            if (kind_is_va(kind)) {
                e.prm = prm_va;
                prm_va += KIND_PK[kind];
            } else {
                e.prm = prm;
                prm += KIND_PK[kind];
            }
            e.lim = lim;
            if (KIND_LIMW[kind]) lim += nu;
            const uint32_t *nd = &T.nodes[kind][(size_t)id * nu];
            const uint32_t *sl = &T.slots[kind][(size_t)id * T.kslots[kind]];
            H.node_off.push_back((uint32_t)H.nodes_flat.size());
            H.slot_off.push_back((uint32_t)H.slots_flat.size());
            H.nodes_flat.insert(H.nodes_flat.end(), nd, nd + nu);
            H.slots_flat.insert(H.slots_flat.end(), sl, sl + T.kslots[kind]);
            uint32_t kp = 0, kg = 0;
            for (int ru = 0; ru < nu; ru++) {
                if ((KIND_RMASK[kind] >> ru & 1) && nd[ru] != 0) rows[nd[ru]].push_back(e.stage_res + ru);
                const unsigned gm = kind_gmask(kind, ru), pm = kind_pmask(kind, ru);
                for (int cu = 0; cu < nu; cu++) {
                    if (!(pm >> cu & 1)) continue;
                    const uint32_t slot = sl[kp++];
                    if (gm >> cu & 1) {
                        if (slot != T.nnz) slots[slot].push_back(e.stage_jac + kg);
                        kg++;
                    }
                }
            }
            stage += kg;
            // charge cells: q per row with a C entry, then one per C entry,
            // added in the serial order the transient evaluation uses
            bool charged = false;
            for (int ru = 0; ru < nu; ru++) charged = charged || kind_cmask(kind, ru) != 0;
            e.stage_q = ESP_NONE;
            e.stage_c = ESP_NONE;
            e.tape = KIND_NSITE[kind] > 0 ? tape : ESP_NONE;
            tape += (uint32_t)KIND_NSITE[kind];
            if (charged) {
                e.stage_q = stage;
                stage += nu;
                e.stage_c = stage;
                uint32_t kc = 0;
                kp = 0;
                for (int ru = 0; ru < nu; ru++) {
                    const unsigned cm = kind_cmask(kind, ru), pm = kind_pmask(kind, ru);
                    if (cm != 0 && nd[ru] != 0) qrows[nd[ru]].push_back(e.stage_q + ru);
                    for (int cu = 0; cu < nu; cu++) {
                        if (!(pm >> cu & 1)) continue;
                        const uint32_t slot = sl[kp++];
                        if (cm >> cu & 1) {
                            if (slot != T.nnz) cslots[slot].push_back(e.stage_c + kc);
                            kc++;
                        }
                    }
                }
                stage += kc;
            }
            inst.push_back(e);
        }
    }
    H.n_tape = tape;
    H.n_inst = (uint32_t)inst.size();
    H.n_stage = stage;
    H.n_prm = prm_va;
    H.n_prm_lds = prm_lds;
    H.n_lim = lim;
    H.inst_bytes.resize(inst.size() * sizeof(InstEntry));
    if (!inst.empty()) std::memcpy(H.inst_bytes.data(), inst.data(), H.inst_bytes.size());
    H.rptr.assign(1, 0);
    H.ridx.clear();
    for (uint32_t r = 0; r < T.n; r++) {
        H.ridx.insert(H.ridx.end(), rows[r].begin(), rows[r].end());
        H.rptr.push_back((uint32_t)H.ridx.size());
    }
    H.gptr.assign(1, 0);
    H.gidx.clear();
    for (uint32_t s = 0; s < T.nnz; s++) {
        H.gidx.insert(H.gidx.end(), slots[s].begin(), slots[s].end());
        H.gptr.push_back((uint32_t)H.gidx.size());
    }
    H.qptr.assign(1, 0);
    H.qidx.clear();
    for (uint32_t r = 0; r < T.n; r++) {
        H.qidx.insert(H.qidx.end(), qrows[r].begin(), qrows[r].end());
        H.qptr.push_back((uint32_t)H.qidx.size());
    }
    H.cptr.assign(1, 0);
    H.cidx.clear();
    for (uint32_t s = 0; s < T.nnz; s++) {
        H.cidx.insert(H.cidx.end(), cslots[s].begin(), cslots[s].end());
        H.cptr.push_back((uint32_t)H.cidx.size());
    }
    // the LDS blob: every table at a word offset
    WaveBlob &B = H.at;
    std::vector<uint32_t> &w = H.blob;
    w.clear();
    auto put = [&w](const uint32_t *p, size_t k) {
        const uint32_t at = (uint32_t)w.size();
        w.insert(w.end(), p, p + k);
        return at;
    };
    B.n = T.n;
    B.nnz = T.nnz;
    B.cap = cap;
    B.n_inst = H.n_inst;
    B.n_stage = H.n_stage;
    B.n_prm = H.n_prm;
    B.n_prm_lds = H.n_prm_lds;
    B.n_lim = H.n_lim;
    B.tri = T.tridiagonal ? 1u : 0u;
    B.diag_fill = T.diag_fill;
    B.col_ptr = put(T.col_ptr.data(), T.col_ptr.size());
    B.row_idx = put(T.row_idx.data(), T.row_idx.size());
    B.q = put(T.q.data(), T.q.size());
    B.tri_pos = put(T.tri_pos.data(), T.tridiagonal ? T.tri_pos.size() : 0);
    B.diag = put(T.diag.data(), T.diag.size());
    {
        std::vector<uint32_t> packed((T.n + 3) / 4, 0);
        if (T.n) std::memcpy(packed.data(), T.cur_row.data(), T.n);
        B.cur_row = put(packed.data(), packed.size());
    }
    {
        std::vector<uint32_t> packed(H.inst_bytes.size() / 4);
        if (!packed.empty()) std::memcpy(packed.data(), H.inst_bytes.data(), H.inst_bytes.size());
        B.inst = put(packed.data(), packed.size());
    }
    B.nodes = put(H.nodes_flat.data(), H.nodes_flat.size());
    B.slots = put(H.slots_flat.data(), H.slots_flat.size());
    B.node_off = put(H.node_off.data(), H.node_off.size());
    B.slot_off = put(H.slot_off.data(), H.slot_off.size());
    B.rptr = put(H.rptr.data(), H.rptr.size());
    B.ridx = put(H.ridx.data(), H.ridx.size());
    B.gptr = put(H.gptr.data(), H.gptr.size());
    B.gidx = put(H.gidx.data(), H.gidx.size());
    B.qptr = put(H.qptr.data(), H.qptr.size());
    B.qidx = put(H.qidx.data(), H.qidx.size());
    B.cptr = put(H.cptr.data(), H.cptr.size());
    B.cidx = put(H.cidx.data(), H.cidx.size());
    B.n_tape = H.n_tape;
    if (w.size() & 1) w.push_back(0);  // the lane block after it is 8-byte aligned
    B.words = (uint32_t)w.size();
}

}  // namespace esp
