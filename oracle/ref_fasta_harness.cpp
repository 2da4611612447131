/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * Compiles the reference's own FASTA parser (src/io/loader/parser/fasta.cpp),
 * with src/encoder.cpp and src/database.cpp, from its sources under the
 * reference tree with its production flags, and exposes one C entry point that
 * loads a file and returns the pairwise::database layout (the records' blocks
 * back to back, offset and size in blocks per record; src/pairwise/
 * database.cu:24-72 packs them the same way) plus the descriptions.  The build
 * recipe is oracle/Makefile (target `ref`); the output goes to oracle/_ref/,
 * which is never committed.  Tests use it to pin oracle/fasta_oracle.c and the
 * GPU loader; bench.py may time it as the CPU baseline.
 *
 * Nothing from the reference is copied here: the translation unit is pulled
 * in by #include from the reference tree at build time.
 */
#include "io/loader/parser/fasta.cpp"

#include <cstdint>
#include <cstring>
#include <string>

/*
 * Loads `filename` with museqa::io::parser::fasta.  Returns the number of
 * records (or -1 on error); *nblocks and *desc_bytes receive the totals.
 * When the output pointers are non-null they receive the layout: blocks,
 * offset / size per record (blocks), and the descriptions back to back with
 * desc_offset per record (count + 1 entries).  Capacities come from a first
 * call with null outputs.
 */
extern "C" __attribute__((visibility("default")))
int64_t ref_fasta_load(const char* filename, uint16_t* blocks, int64_t* offset, int32_t* size, char* desc,
                       int64_t* desc_offset, int64_t* nblocks, int64_t* desc_bytes)
{
    try {
        const auto db = museqa::io::parser::fasta(filename);
        int64_t at = 0, dat = 0;
        for (size_t i = 0; i < db.count(); ++i) {
            const auto& e = db[i];
            const auto n = (int64_t) e.contents.size();
            if (offset) offset[i] = at;
            if (size) size[i] = (int32_t) n;
            if (blocks)
                for (int64_t b = 0; b < n; ++b) blocks[at + b] = e.contents.block(b);
            if (desc_offset) desc_offset[i] = dat;
            if (desc) std::memcpy(desc + dat, e.description.data(), e.description.size());
            at += n;
            dat += (int64_t) e.description.size();
        }
        if (desc_offset) desc_offset[db.count()] = dat;
        if (nblocks) *nblocks = at;
        if (desc_bytes) *desc_bytes = dat;
        return (int64_t) db.count();
    } catch (...) {
        return -1;
    }
}
