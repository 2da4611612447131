/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * A plain-C restatement of the reference's FASTA loading path over an
 * in-memory text, used only by tests/ and by bench.py's cpu_baseline leg as
 * the checker.  Nothing in msa_amd/ links, loads or calls it.
 *
 * What it follows:
 *   - io::parser::fasta / extract_to_db (src/io/loader/parser/fasta.cpp:26-70)
 *     with std::getline / peek semantics: lines end at '\n' (a '\r' stays in
 *     the line); a final piece without '\n' is a line, an empty final piece is
 *     not; lines are skipped until one starts with '>'; its rest is the
 *     description; the following lines are the contents while the next byte is
 *     not '>' and the line is not empty; an empty line ends the record and
 *     the lines after it are skipped up to the next '>' line;
 *   - encoder::encode (src/encoder.cpp:76-98): codes by letter (case-folded),
 *     '-' the gap code 0x19, anything else the end code 0x18; three codes per
 *     uint16 block at bits 1, 6 and 11; bit 0 set on the last block when the
 *     length is not a multiple of three, whose unused slots hold 0x18;
 *   - pairwise::database (src/pairwise/database.cu:24-72): the records'
 *     blocks back to back; offset and size per record, in blocks.
 * Descriptions are reported as [start, end) byte ranges of the text.
 */
#include <stdint.h>
#include <string.h>

/* src/encoder.cpp:31-34, indexed by upper - 'A' */
static const uint8_t kLetter[26] = {0x00, 0x14, 0x01, 0x06, 0x08, 0x0E, 0x03, 0x09, 0x0A, 0x15, 0x0C, 0x0B, 0x0D,
                                    0x05, 0x17, 0x0F, 0x07, 0x04, 0x10, 0x02, 0x17, 0x13, 0x11, 0x17, 0x12, 0x16};

static uint8_t encode_byte(uint8_t c)
{
    const uint8_t upper = (c >= 'a' && c <= 'z') ? (uint8_t)(c - 'a' + 'A') : c;
    if (upper >= 'A' && upper <= 'Z') return kLetter[upper - 'A'];
    return c == '-' ? 0x19 : 0x18;
}

/* One getline: the line at *pos, its end (exclusive of '\n'); *pos moves past
 * the '\n'.  Returns 0 when nothing is left (getline fails). */
static int next_line(const uint8_t* text, int64_t n, int64_t* pos, int64_t* start, int64_t* end)
{
    if (*pos >= n) return 0;
    const uint8_t* nl = memchr(text + *pos, '\n', (size_t)(n - *pos));
    *start = *pos;
    *end = nl ? (int64_t)(nl - text) : n;
    *pos = nl ? *end + 1 : n;
    return 1;
}

/* Parses the text.  With null outputs it only counts; otherwise it fills
 * blocks / offset / size / desc_start / desc_end (capacities from a counting
 * call).  Returns the number of records; *total_blocks gets the block count. */
int64_t fasta_oracle(const uint8_t* text, int64_t n, uint16_t* blocks, int64_t* offset, int32_t* size,
                     int64_t* desc_start, int64_t* desc_end, int64_t* total_blocks)
{
    int64_t pos = 0, records = 0, nblocks = 0;
    for (;;) {
        /* skip to a '>' line */
        int64_t s = 0, e = 0;
        int found = 0;
        while (next_line(text, n, &pos, &s, &e))
            if (e > s && text[s] == '>') { found = 1; break; }
        if (!found) break;
        if (desc_start) desc_start[records] = s + 1, desc_end[records] = e;
        if (offset) offset[records] = nblocks;
        /* contents: while the next byte is not '>' and the line is not empty */
        int64_t length = 0;
        uint16_t cur = 0;
        while (pos < n && text[pos] != '>') {
            int64_t ls, le;
            next_line(text, n, &pos, &ls, &le);
            if (le == ls) break;
            for (int64_t p = ls; p < le; ++p, ++length) {
                const int slot = (int)(length % 3);
                cur |= (uint16_t)(encode_byte(text[p]) << (1 + 5 * slot));
                if (slot == 2) {
                    if (blocks) blocks[nblocks] = cur;
                    ++nblocks;
                    cur = 0;
                }
            }
        }
        if (length % 3) {
            for (int slot = (int)(length % 3); slot < 3; ++slot) cur |= (uint16_t)(0x18 << (1 + 5 * slot));
            if (blocks) blocks[nblocks] = (uint16_t)(cur | 1);
            ++nblocks;
        }
        if (size) size[records] = (int32_t)((length + 2) / 3);
        ++records;
    }
    if (total_blocks) *total_blocks = nblocks;
    return records;
}
