// Circuit templates of the lane engine (host): node numbering, the frozen
// MNA pattern, slot tapes, the shared column order. See engine.h Template.
#include <algorithm>
#include <cstring>

#include "engine.h"

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
    for (int p = 0; p < KIND_NPORT[kind]; p++) nodes[kind].push_back(ports[p]);
    for (int u = KIND_NPORT[kind]; u < KIND_NU[kind]; u++) {
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

}  // namespace esp

