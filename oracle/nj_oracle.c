/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * A plain-C restatement of the reference's sequential neighbor-joining
 * (src/phylogeny/njoining/impl/sequential.cpp), used only by tests/ and the
 * cpu_baseline leg of bench.py as the checker and the CPU baseline.  Nothing in
 * msa_amd/ links, loads or calls it.
 *
 * What it follows (behaviour, float for float):
 *   - input: the pairwise distance_matrix, float, packed in pair order
 *     c = i(i-1)/2 + j (src/pairwise/pairwise.cuh:99-107), 0 on the diagonal,
 *     copied into a symmetric matrix (src/phylogeny/matrix.cuh:124-131);
 *   - cache_init (sequential.cpp:79-94): row sums accumulated in float, row k
 *     in index order; the reference zeroes cache[0..count-2] only and reads
 *     cache[count-1] uninitialised (new float[n], src/allocator.hpp:110), which
 *     this restatement takes as 0;
 *   - q_transform (:123-127): Q = (count-2) * d(i,j) - S_i - S_j in float,
 *     left to right;
 *   - pick_joinable (:159-177): pairs (i, j), i = 1.., j < i, in that order;
 *     the first strictly greater Q wins, starting from -infinity;
 *   - raise_candidate (:136-146): dx = .5 * d (double) + (S_x - S_y) /
 *     (2 * (count-2)) (float division), rounded to float; dy = d - dx; with
 *     x = the larger index i, y = j;
 *   - join_pair (:228-262): for every OTU i (x and y included) previous =
 *     d(i,x) + d(i,y), updated = (previous - d(x,y)) * .5 (double),
 *     S_i += updated - previous (double, rounded to float), new_sum += updated
 *     (double, rounded to float) in index order; row/column x <- updated;
 *     S_x <- new_sum; the new OTU takes x's place;
 *   - update_cache (:188-221): symmetric variant: y swaps with index 0 and
 *     index 0 is removed (every index shifts down by one); linear variant: y
 *     swaps with the last index, which is removed;
 *   - star::join (src/phylogeny/njoining/star.cuh:63-80): the parent gets
 *     children ordered by node id, each child its branch length and parent,
 *     level = max(children's levels) + 1; parents are numbered n, n+1, ...;
 *     the tree has 2n-1 nodes, the root is node 2n-2.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NJ_UNDEFINED 0xFFFFFFFFu

typedef struct {
    uint32_t parent;
    uint32_t child[2];
    float distance;      /* branch length to the parent; FLT_MAX for the root */
    uint32_t level;
} nj_node;

/* variant 0: symmetric matrix (the reference's default), 1: linear.
 * `mistake` plants a plausible error for the tests that show the pins catch it:
 * 1 = Q with (count - 1), 2 = pick the smallest Q, 3 = update without the .5,
 * 4 = the other variant's removal; 0 = none. */
static int nj_core(const float* dist, int32_t n, int32_t variant, int32_t mistake, nj_node* out)
{
    if (n < 2) return 0;
    float* m = (float*)malloc(sizeof(float) * (size_t)n * (size_t)n);
    int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);     /* logical index -> matrix row */
    uint32_t* map = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);   /* logical index -> node id */
    float* cache = (float*)malloc(sizeof(float) * (size_t)n);
    float* updated_row = (float*)malloc(sizeof(float) * (size_t)n);
    if (!m || !slot || !map || !cache || !updated_row) {
        free(m), free(slot), free(map), free(cache), free(updated_row);
        return -1;
    }
    for (int32_t i = 0; i < n; ++i) {
        m[(size_t)i * n + i] = 0.0f;
        for (int32_t j = 0; j < i; ++j) {
            const float d = dist[(size_t)i * (i - 1) / 2 + j];
            m[(size_t)i * n + j] = d;
            m[(size_t)j * n + i] = d;
        }
    }
    for (int32_t i = 0; i < 2 * n - 1; ++i) {
        out[i].parent = NJ_UNDEFINED;
        out[i].child[0] = out[i].child[1] = NJ_UNDEFINED;
        out[i].distance = 3.40282347e+38f;
        out[i].level = 0;
    }
    size_t count = (size_t)n;
    int32_t* base = slot;              /* the symmetric variant removes index 0 by moving this */
    uint32_t* mbase = map;
    float* cbase = cache;
    for (int32_t i = 0; i < n; ++i) slot[i] = i, map[i] = (uint32_t)i, cache[i] = 0.0f;
    for (size_t i = 0; i + 1 < count; ++i)
        for (size_t j = i + 1; j < count; ++j) {
            const float cur = m[(size_t)i * n + j];
            cache[i] += cur;
            cache[j] += cur;
        }

#define M(a, b) m[(size_t)base[a] * (size_t)n + (size_t)base[b]]
    uint32_t parent = (uint32_t)n;
    while (count > 1) {
        /* pick_joinable over all pairs */
        size_t bx = NJ_UNDEFINED, by = NJ_UNDEFINED;
        float best = -INFINITY;
        for (size_t i = 1; i < count; ++i)
            for (size_t j = 0; j < i; ++j) {
                const float q = (float)(count - (mistake == 1 ? 1 : 2)) * M(i, j) - cbase[i] - cbase[j];
                if (mistake == 2 ? (q < -best || best == -INFINITY) : q > best)
                    best = mistake == 2 ? -q : q, bx = i, by = j;
            }
        if (bx == NJ_UNDEFINED) {         /* no finite Q: the reference's pick is undefined */
            free(m), free(slot), free(map), free(cache), free(updated_row);
            return -2;
        }
        /* raise_candidate */
        const float dxy = M(bx, by);
        const float pairsum = cbase[bx] - cbase[by];
        const float dx = (float)(.5 * (double)dxy + (double)(pairsum / (float)(2 * (count - 2))));
        const float dy = dxy - dx;

        /* star::join */
        {
            const uint32_t a = mbase[bx], b = mbase[by];
            const uint32_t c1 = a < b ? a : b, c2 = a < b ? b : a;
            const float d1 = a < b ? dx : dy, d2 = a < b ? dy : dx;
            out[c1].distance = d1, out[c1].parent = parent;
            out[c2].distance = d2, out[c2].parent = parent;
            out[parent].level = (out[c1].level > out[c2].level ? out[c1].level : out[c2].level) + 1;
            out[parent].child[0] = c1, out[parent].child[1] = c2;
        }

        /* join_pair */
        float new_sum = 0.0f;
        for (size_t i = 0; i < count; ++i) {
            const float previous = M(i, bx) + M(i, by);
            const double updated = (double)(previous - dxy) * (mistake == 3 ? 1.0 : .5);
            updated_row[i] = (float)updated;
            cbase[i] = (float)((double)cbase[i] + (updated - (double)previous));
            new_sum = (float)((double)new_sum + updated);
        }
        for (size_t i = 0; i < count; ++i) M(i, bx) = M(bx, i) = updated_row[i];
        cbase[bx] = new_sum;
        mbase[bx] = parent++;

        /* update_cache */
        if ((variant == 0) != (mistake == 4)) {
            if (by != 0) {
                float tc = cbase[0]; cbase[0] = cbase[by]; cbase[by] = tc;
                int32_t ts = base[0]; base[0] = base[by]; base[by] = ts;
                uint32_t tm = mbase[0]; mbase[0] = mbase[by]; mbase[by] = tm;
            }
            ++base, ++mbase, ++cbase;
        } else {
            const size_t last = count - 1;
            if (by != last) {
                float tc = cbase[last]; cbase[last] = cbase[by]; cbase[by] = tc;
                int32_t ts = base[last]; base[last] = base[by]; base[by] = ts;
                uint32_t tm = mbase[last]; mbase[last] = mbase[by]; mbase[by] = tm;
            }
        }
        --count;
    }
#undef M
    free(m), free(slot), free(map), free(cache), free(updated_row);
    return 0;
}

int nj_oracle(const float* dist, int32_t n, int32_t variant, nj_node* out)
{
    return nj_core(dist, n, variant, 0, out);
}

int nj_oracle_planted(const float* dist, int32_t n, int32_t variant, int32_t mistake, nj_node* out)
{
    return nj_core(dist, n, variant, mistake, out);
}
