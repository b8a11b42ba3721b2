// Thin C entry points over the reference's own inv_list (src/genie/table/
// inv_list.cc), which is the only piece of the reference's match path that
// compiles from its own sources with the tools here (everything else needs
// CUDA, Thrust or Boost).  Built by oracle/Makefile into oracle/_ref/ and used
// only by tests/ to pin the oracle's inversion against the reference code.
#include <vector>
#include <cstdint>
#include "genie/table/inv_list.h"

extern "C" {

// inv_list::invert on one column; writes lo/hi and, for every value lo..hi,
// the list length into lens (caller sizes it) and the ids into ids.
int ref_invert(const int32_t* col, int32_t n, int32_t* lo, int32_t* hi,
               int32_t* lens, int64_t lens_cap, int32_t* ids, int64_t ids_cap)
{
    std::vector<int> v(col, col + n);
    genie::table::inv_list L;
    L.invert(v);
    *lo = L.min();
    *hi = L.max();
    int64_t nv = (int64_t)L.max() - L.min() + 1, p = 0;
    if (nv > lens_cap) return -1;
    for (int64_t j = 0; j < nv; ++j) {
        std::vector<int>* lst = L.index((int)(L.min() + j));
        lens[j] = lst ? (int32_t)lst->size() : 0;
        for (size_t t = 0; lst && t < lst->size(); ++t) {
            if (p >= ids_cap) return -1;
            ids[p++] = (*lst)[t];
        }
    }
    return 0;
}

// inv_list::invert_bijectMap on CSR rows.
int ref_invert_biject(const int64_t* row_ptr, const int32_t* vals, int32_t n_rows,
                      int32_t* lo, int32_t* hi, int32_t* lens, int64_t lens_cap,
                      int32_t* ids, int64_t ids_cap)
{
    std::vector<std::vector<int> > rows(n_rows);
    for (int32_t i = 0; i < n_rows; ++i)
        rows[i].assign(vals + row_ptr[i], vals + row_ptr[i + 1]);
    genie::table::inv_list L;
    L.invert_bijectMap(rows);
    *lo = L.min();
    *hi = L.max();
    int64_t nv = (int64_t)L.max() - L.min() + 1, p = 0;
    if (nv > lens_cap) return -1;
    for (int64_t j = 0; j < nv; ++j) {
        std::vector<int>* lst = L.index((int)(L.min() + j));
        lens[j] = lst ? (int32_t)lst->size() : 0;
        for (size_t t = 0; lst && t < lst->size(); ++t) {
            if (p >= ids_cap) return -1;
            ids[p++] = (*lst)[t];
        }
    }
    return 0;
}

}
