/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * A plain-C statement of the progressive profile alignment that the GPU
 * pgalign stage (msa_amd/csrc/pgalign.hip) computes, used only by tests/ as
 * the checker.  Nothing in msa_amd/ links, loads or calls it.
 *
 * The reference's profile aligner is a stub: its "myers" algorithm returns an
 * empty alignment (src/pgalign/myers/impl/sequential.cpp:29-32), so there is
 * no reference output to restate.  What is kept from the reference:
 *   - the sequence a row holds: its first unpadded() codes, where unpadded()
 *     is the padded length minus the padding codes at positions 1 and 2 of
 *     the last block (src/sequence.hpp:110-119);
 *   - the gapped row: the column of every residue, ascending, and the
 *     alignment's length (src/pgalign/sequence.cuh: m_columns, length());
 *   - the guide tree from neighbor joining: join t is node n + t with two
 *     children, smaller id first (src/phylogeny/tree.cuh, star.cuh:63-80).
 * The objective is the one the reference's pipeline implies: every join of the
 * tree, children before parents, aligns the two children's alignments
 * (profiles) optimally under the sum-of-pairs score with the pairwise stage's
 * table T and linear gap penalty p: sigma(x, y) = T[x][y] for residues,
 * sigma(residue, gap) = sigma(gap, residue) = -p, sigma(gap, gap) = 0.  For
 * two single sequences that is Needleman-Wunsch with T and p, the score of
 * src/pairwise/needleman/impl/sequential.cpp:29-68 on clean input.
 *
 * Join of A (rows, nA members, LA columns) and B (columns, nB members, LB):
 *   S(i, j) = sum_x sum_y cA[i][x] cB[j][y] sigma(x, y)   (x, y over 0..24 and gap)
 *   dA[i]   = -p * rA[i] * nB      (column i of A against a gap column)
 *   dB[j]   = -p * rB[j] * nA
 *   H[0][0] = 0, H[i][0] = H[i-1][0] + dA[i], H[0][j] = H[0][j-1] + dB[j],
 *   Ed = H[i-1][j-1] + S(i, j), Eu = H[i-1][j] + dA[i], L = H[i][j-1] + dB[j],
 *   H[i][j] = max(Ed, Eu, L), ties resolved diagonal, then up, then left;
 * with c the per-column code counts, r the residues and 64-bit scores.  The
 * traceback from (LA, LB) follows the recorded choice (up along column 0,
 * left along row 0) and the new alignment's columns are the path's steps.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PG_PADDING 0x18
#define PG_CODES 26            /* 25 residue codes and the gap */
#define PG_GAP 25
#define PG_NONE 0xFFFFFFFFu

typedef struct {
    uint32_t parent;
    uint32_t child[2];
    float distance;
    uint32_t level;
} pg_node;

static inline unsigned pg_access(const uint16_t* seq, size_t offset)
{
    static const unsigned shift[3] = {1, 6, 11};
    return (unsigned)(seq[offset / 3] >> shift[offset % 3]) & 0x1Fu;
}

/* sequence::unpadded (src/sequence.hpp:110-119). */
int64_t pg_unpadded(const uint16_t* seq, int32_t blocks)
{
    if (blocks <= 0) return 0;
    int64_t length = 3 * (int64_t)blocks;
    for (int k = 1; k < 3; ++k) length -= pg_access(seq, 3 * (size_t)(blocks - 1) + k) == PG_PADDING;
    return length;
}

/* One join: A's members' columns are colA (residues of A in member order), B's
 * colB; returns the new length and rewrites both column arrays; *score gets
 * the join's optimal sum-of-pairs score.  mistake != 0 plants an error (tests
 * of the pins only): 1 = sigma(gap, gap) = -p, 2 = a gap column costs -p per
 * member instead of per residue, 3 = ties resolved left before diagonal. */
static int64_t pg_join(const uint8_t* codesA, int32_t* colA, int64_t resA, int64_t nA, int64_t LA,
                       const uint8_t* codesB, int32_t* colB, int64_t resB, int64_t nB, int64_t LB,
                       const int32_t* table625, int32_t p, int mistake, int64_t* score)
{
    int64_t* cA = calloc((size_t)(LA ? LA : 1) * PG_CODES, sizeof(int64_t));
    int64_t* cB = calloc((size_t)(LB ? LB : 1) * PG_CODES, sizeof(int64_t));
    for (int64_t r = 0; r < resA; ++r) cA[colA[r] * PG_CODES + codesA[r]] += 1;
    for (int64_t r = 0; r < resB; ++r) cB[colB[r] * PG_CODES + codesB[r]] += 1;
    int64_t sigma[PG_CODES][PG_CODES];
    for (int x = 0; x < PG_CODES; ++x)
        for (int y = 0; y < PG_CODES; ++y) {
            if (x < 25 && y < 25) sigma[x][y] = table625[x * 25 + y];
            else if (x == PG_GAP && y == PG_GAP) sigma[x][y] = mistake == 1 ? -p : 0;
            else sigma[x][y] = -p;
        }
    int64_t* rA = calloc((size_t)(LA ? LA : 1), sizeof(int64_t));
    int64_t* rB = calloc((size_t)(LB ? LB : 1), sizeof(int64_t));
    for (int64_t i = 0; i < LA; ++i) {
        int64_t r = 0;
        for (int x = 0; x < 25; ++x) r += cA[i * PG_CODES + x];
        rA[i] = r;
        cA[i * PG_CODES + PG_GAP] = nA - r;
    }
    for (int64_t j = 0; j < LB; ++j) {
        int64_t r = 0;
        for (int y = 0; y < 25; ++y) r += cB[j * PG_CODES + y];
        rB[j] = r;
        cB[j * PG_CODES + PG_GAP] = nB - r;
    }
    /* V[i][y] = sum_x cA[i][x] sigma(x, y) */
    int64_t* V = calloc((size_t)(LA ? LA : 1) * PG_CODES, sizeof(int64_t));
    for (int64_t i = 0; i < LA; ++i)
        for (int y = 0; y < PG_CODES; ++y) {
            int64_t v = 0;
            for (int x = 0; x < PG_CODES; ++x) v += cA[i * PG_CODES + x] * sigma[x][y];
            V[i * PG_CODES + y] = v;
        }
    uint8_t* dir = malloc((size_t)(LA + 1) * (size_t)(LB + 1));
    int64_t* prev = malloc(sizeof(int64_t) * (size_t)(LB + 1));
    int64_t* cur = malloc(sizeof(int64_t) * (size_t)(LB + 1));
    prev[0] = 0;
    for (int64_t j = 1; j <= LB; ++j) {
        const int64_t dB = mistake == 2 ? -p * nA : -p * rB[j - 1] * nA;
        prev[j] = prev[j - 1] + dB;
        dir[j] = 2;
    }
    for (int64_t i = 1; i <= LA; ++i) {
        const int64_t dA = mistake == 2 ? -p * nB : -p * rA[i - 1] * nB;
        cur[0] = prev[0] + dA;
        dir[i * (LB + 1)] = 1;
        for (int64_t j = 1; j <= LB; ++j) {
            int64_t s = 0;
            for (int y = 0; y < PG_CODES; ++y) s += V[(i - 1) * PG_CODES + y] * cB[(j - 1) * PG_CODES + y];
            const int64_t dB = mistake == 2 ? -p * nA : -p * rB[j - 1] * nA;
            const int64_t ed = prev[j - 1] + s, eu = prev[j] + dA, l = cur[j - 1] + dB;
            int64_t h;
            uint8_t d;
            if (mistake == 3) {
                if (l >= ed && l >= eu) h = l, d = 2;
                else if (ed >= eu) h = ed, d = 0;
                else h = eu, d = 1;
            } else {
                const int64_t e = ed >= eu ? ed : eu;
                const uint8_t de = ed >= eu ? 0 : 1;
                if (e >= l) h = e, d = de;
                else h = l, d = 2;
            }
            cur[j] = h;
            dir[i * (LB + 1) + j] = d;
        }
        int64_t* t = prev;
        prev = cur;
        cur = t;
    }
    *score = prev[LB];
    /* Traceback: steps counted from the end. */
    int32_t* mapA = malloc(sizeof(int32_t) * (size_t)(LA ? LA : 1));
    int32_t* mapB = malloc(sizeof(int32_t) * (size_t)(LB ? LB : 1));
    int64_t i = LA, j = LB, t = 0;
    while (i > 0 || j > 0) {
        const uint8_t d = i == 0 ? 2 : (j == 0 ? 1 : dir[i * (LB + 1) + j]);
        if (d == 0) mapA[--i] = (int32_t)t, mapB[--j] = (int32_t)t;
        else if (d == 1) mapA[--i] = (int32_t)t;
        else mapB[--j] = (int32_t)t;
        ++t;
    }
    for (int64_t r = 0; r < resA; ++r) colA[r] = (int32_t)(t - 1 - mapA[colA[r]]);
    for (int64_t r = 0; r < resB; ++r) colB[r] = (int32_t)(t - 1 - mapB[colB[r]]);
    free(cA); free(cB); free(rA); free(rB); free(V); free(dir); free(prev); free(cur); free(mapA); free(mapB);
    return t;
}

/*
 * Progressive alignment of n sequences along a guide tree of 2n-1 nodes
 * (leaves 0..n-1, join t = node n+t, children smaller id first).  Output:
 * columns[] holds, for every sequence in database order, the column of each
 * of its unpadded() residues; *length the alignment's length.  join_scores,
 * when not null, receives each join's optimal sum-of-pairs score (n-1).
 * Returns 0, or -1 for an invalid tree or a gap code (0x19) in a residue.
 */
int pg_oracle_planted(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t n,
                      const pg_node* nodes, const int32_t* table625, int32_t penalty, int mistake,
                      int32_t* columns, int64_t* length, int64_t* join_scores)
{
    *length = 0;
    if (n <= 0) return 0;
    int64_t* len = malloc(sizeof(int64_t) * (size_t)n);
    int64_t* out_at = malloc(sizeof(int64_t) * (size_t)n);
    int64_t total = 0;
    for (int32_t s = 0; s < n; ++s) {
        len[s] = pg_unpadded(blocks + offset[s], size[s]);
        out_at[s] = total;
        total += len[s];
        for (int64_t r = 0; r < len[s]; ++r)
            if (pg_access(blocks + offset[s], (size_t)r) > PG_PADDING) {
                free(len); free(out_at);
                return -1;
            }
    }
    const int32_t nodes_n = 2 * n - 1;
    /* Leaf order: members(node) = members(child0) ++ members(child1), built by
     * an explicit post-order walk from the root. */
    int32_t* lo = malloc(sizeof(int32_t) * (size_t)nodes_n);
    int32_t* hi = malloc(sizeof(int32_t) * (size_t)nodes_n);
    int32_t* order = malloc(sizeof(int32_t) * (size_t)n);
    int32_t* stack = malloc(sizeof(int32_t) * (size_t)(2 * nodes_n + 2));
    int32_t sp = 0, placed = 0;
    stack[sp++] = nodes_n - 1;
    while (sp) {
        const int32_t v = stack[--sp];
        if (v >= 0) {
            if (v < n) {
                lo[v] = placed;
                order[placed++] = v;
                hi[v] = placed;
                continue;
            }
            stack[sp++] = -v - 1;     /* revisit after the children */
            const uint32_t c0 = nodes[v].child[0], c1 = nodes[v].child[1];
            if (c0 >= (uint32_t)v || c1 >= (uint32_t)v || c0 == c1) {
                free(len); free(out_at); free(lo); free(hi); free(order); free(stack);
                return -1;
            }
            stack[sp++] = (int32_t)c1;
            stack[sp++] = (int32_t)c0;
        } else {
            const int32_t u = -v - 1;
            lo[u] = lo[nodes[u].child[0]];
            hi[u] = hi[nodes[u].child[1]];
        }
    }
    if (placed != n) {
        free(len); free(out_at); free(lo); free(hi); free(order); free(stack);
        return -1;
    }
    /* Residues in leaf order: codes and current columns. */
    int64_t* res_at = malloc(sizeof(int64_t) * (size_t)(n + 1));
    res_at[0] = 0;
    for (int32_t k = 0; k < n; ++k) res_at[k + 1] = res_at[k] + len[order[k]];
    uint8_t* codes = malloc((size_t)(total ? total : 1));
    int32_t* col = malloc(sizeof(int32_t) * (size_t)(total ? total : 1));
    for (int32_t k = 0; k < n; ++k) {
        const int32_t s = order[k];
        for (int64_t r = 0; r < len[s]; ++r) {
            codes[res_at[k] + r] = (uint8_t)pg_access(blocks + offset[s], (size_t)r);
            col[res_at[k] + r] = (int32_t)r;
        }
    }
    int64_t* plen = malloc(sizeof(int64_t) * (size_t)nodes_n);
    for (int32_t s = 0; s < n; ++s) plen[s] = len[s];
    for (int32_t v = n; v < nodes_n; ++v) {
        const uint32_t a = nodes[v].child[0], b = nodes[v].child[1];
        const int64_t ra0 = res_at[lo[a]], ra1 = res_at[hi[a]], rb0 = res_at[lo[b]], rb1 = res_at[hi[b]];
        int64_t sc = 0;
        plen[v] = pg_join(codes + ra0, col + ra0, ra1 - ra0, hi[a] - lo[a], plen[a], codes + rb0, col + rb0, rb1 - rb0,
                          hi[b] - lo[b], plen[b], table625, penalty, mistake, &sc);
        if (join_scores) join_scores[v - n] = sc;
    }
    *length = plen[nodes_n - 1];
    for (int32_t k = 0; k < n; ++k) {
        const int32_t s = order[k];
        memcpy(columns + out_at[s], col + res_at[k], sizeof(int32_t) * (size_t)len[s]);
    }
    free(len); free(out_at); free(lo); free(hi); free(order); free(stack); free(res_at); free(codes); free(col);
    free(plen);
    return 0;
}

int pg_oracle(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t n, const pg_node* nodes,
              const int32_t* table625, int32_t penalty, int32_t* columns, int64_t* length, int64_t* join_scores)
{
    return pg_oracle_planted(blocks, offset, size, n, nodes, table625, penalty, 0, columns, length, join_scores);
}
