// C++ host layer over libmsa_hip.so: encoder, sequences, databases, FASTA
// loading, scoring tables, distance matrix, the pairwise algorithm registry
// and run().  Behaviour follows the reference files cited in museqa.hpp; the
// alignment itself always runs in the HIP kernels behind the C ABI.
#include "museqa.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <utility>

#include "msa_hip.h"

namespace museqa
{
    namespace
    {
        // Letter -> code, indexed by letter - 'A' (src/encoder.cpp:27-34).
        constexpr encoder::unit kLetterCode[26] = {
            0x00, 0x14, 0x01, 0x06, 0x08, 0x0E, 0x03, 0x09, 0x0A, 0x15, 0x0C, 0x0B, 0x0D,
            0x05, 0x17, 0x0F, 0x07, 0x04, 0x10, 0x02, 0x17, 0x13, 0x11, 0x17, 0x12, 0x16};

        // Code -> letter (src/encoder.cpp:36-43).
        constexpr char kCodeLetter[] = "ACTGRNDQEHILKMFPSWYVBJZX*-";

        [[noreturn]] void raise(const char* fmt, const std::string& arg)
        {
            char buf[512];
            std::snprintf(buf, sizeof(buf), fmt, arg.c_str());
            throw exception(buf);
        }

        void check(int status)
        {
            if (status == MSA_OK) return;
            std::string msg = msa_last_error();
            if (msg.empty()) msg = msa_error_string(status);
            throw exception(msg);
        }
    }

    // ---- encoder ----------------------------------------------------------

    encoder::unit encoder::access(block tgt, uint8_t offset) noexcept
    {
        return (tgt >> (1 + 5 * offset)) & 0x1F;
    }

    encoder::unit encoder::access(const block* tgt, size_t offset) noexcept
    {
        return access(tgt[offset / block_size], offset % block_size);
    }

    encoder::unit encoder::encode(char value) noexcept
    {
        const char upper = (value >= 'a' && value <= 'z') ? char(value - 'a' + 'A') : value;
        if (upper >= 'A' && upper <= 'Z') return kLetterCode[upper - 'A'];
        return value == '-' ? unit(gap) : unit(end);
    }

    encoder::buffer encoder::encode(const char* ptr, size_t size)
    {
        buffer out((size + block_size - 1) / block_size);
        for (size_t i = 0, n = 0; n < size; ++i) {
            // Bit 0 marks the block that holds padding (src/encoder.cpp:86-92).
            block current = (n + block_size > size) ? 1 : 0;
            for (size_t j = 0; j < block_size; ++j, ++n) {
                const unit code = n < size ? encode(ptr[n]) : unit(end);
                current |= block(code << (1 + 5 * j));
            }
            out[i] = current;
        }
        return out;
    }

    char encoder::decode(unit value)
    {
        if (value >= sizeof(kCodeLetter) - 1) throw exception("cannot decode invalid unit");
        return kCodeLetter[value];
    }

    std::string encoder::decode(const block* blocks, size_t nblocks)
    {
        std::string out(nblocks * block_size, '*');
        for (size_t i = 0; i < out.size(); ++i) out[i] = decode(access(blocks, i));
        return out;
    }

    // ---- sequence ---------------------------------------------------------

    encoder::unit sequence::operator[](ptrdiff_t offset) const
    {
        if (offset < 0 || size_t(offset) >= length()) throw std::out_of_range("sequence offset out of range");
        return encoder::access(m_blocks.data(), size_t(offset));
    }

    size_t sequence::unpadded() const noexcept
    {
        if (m_blocks.empty()) return 0;
        size_t len = length();
        for (uint8_t i = 1; i < encoder::block_size; ++i) len -= (encoder::access(m_blocks.back(), i) == padding);
        return len;
    }

    // ---- database ---------------------------------------------------------

    int database::anonymous = 0;

    const database::entry_type& database::operator[](ptrdiff_t offset) const
    {
        if (offset < 0 || size_t(offset) >= count()) throw exception("database offset out of range");
        return m_entries[offset];
    }

    const database::entry_type& database::operator[](const std::string& key) const
    {
        const auto it = m_indeces.find(key);
        if (it == m_indeces.end()) throw exception("cannot find key in database");
        return m_entries[it->second];
    }

    void database::add(const std::string& description, const sequence& elem)
    {
        m_indeces[description] = ptrdiff_t(m_entries.size());
        m_entries.push_back({description, elem});
    }

    void database::add(const sequence& elem)
    {
        char name[32];
        std::snprintf(name, sizeof(name), "anonymous#%d", ++anonymous);
        add(name, elem);
    }

    void database::add(const std::vector<sequence>& elems)
    {
        for (const auto& e : elems) add(e);
    }

    database database::only(const std::set<ptrdiff_t>& keys) const
    {
        database out;
        for (const auto k : keys) {
            const auto& e = (*this)[k];
            out.add(e.description, e.contents);
        }
        return out;
    }

    database database::only(const std::set<std::string>& keys) const
    {
        database out;
        for (const auto& k : keys) {
            const auto& e = (*this)[k];
            out.add(e.description, e.contents);
        }
        return out;
    }

    void database::merge(const database& db)
    {
        const size_t first = m_entries.size();
        m_entries.insert(m_entries.end(), db.m_entries.begin(), db.m_entries.end());
        for (size_t i = first; i < m_entries.size(); ++i) m_indeces[m_entries[i].description] = ptrdiff_t(i);
    }

    database database::load(const std::string& filename)
    {
        // utils::extension (src/utils.hpp:263-267): text after the last '.',
        // or the whole name when there is none.
        const auto dot = filename.find_last_of('.');
        const std::string ext = filename.empty() ? std::string {}
                              : filename.substr(dot == std::string::npos ? 0 : dot + 1);
        if (ext == "fa" || ext == "fasta") return io::parser::fasta(filename);
        raise("unknown database parser '%s'", ext);
    }

    // FASTA (src/io/loader/parser/fasta.cpp:26-70): a record begins at a '>'
    // line and takes the lines that follow up to the next '>' or the first
    // empty line; characters before the first '>' are skipped.
    database io::parser::fasta(const std::string& filename)
    {
        std::fstream file(filename, std::fstream::in);
        if (file.fail()) raise("file does not exist or cannot be read '%s'", filename);
        database result;
        while (!file.eof() && !file.fail()) {
            std::string line, contents;
            bool found = true;
            while (line.empty() || line[0] != '>') {
                if (file.eof()) { found = false; break; }
                std::getline(file, line);
            }
            if (!found) break;
            const std::string description = line.substr(1);
            while (file.peek() != '>' && std::getline(file, line) && !line.empty()) contents.append(line);
            result.add(description, sequence {contents});
        }
        return result;
    }

    database database::load_on_device(const std::string& filename)
    {
        const auto dot = filename.find_last_of('.');
        const std::string ext = filename.empty() ? std::string {}
                              : filename.substr(dot == std::string::npos ? 0 : dot + 1);
        if (ext != "fa" && ext != "fasta") raise("unknown database parser '%s'", ext);
        std::ifstream file(filename, std::ios::binary | std::ios::ate);
        if (!file) raise("file does not exist or cannot be read '%s'", filename);
        std::string text(size_t(file.tellg()), '\0');
        file.seekg(0);
        if (!file.read(&text[0], std::streamsize(text.size()))) raise("cannot read '%s'", filename);
        return io::parser::fasta_on_device(text);
    }

    // The GPU loader builds the sequence store; records become entries in
    // file order, with database::add's key rule.
    database io::parser::fasta_on_device(const std::string& text)
    {
        int devices = 0;
        check(msa_device_count(&devices));
        if (devices < 1) throw exception("no GPU available for the FASTA loader");
        msa_fasta loader = nullptr;
        check(msa_fasta_create(&loader));
        struct closer {
            msa_fasta f;
            ~closer() { msa_fasta_destroy(f); }
        } guard {loader};
        check(msa_fasta_parse(loader, text.data(), int64_t(text.size())));
        msa_fasta_info info {};
        check(msa_fasta_info_get(loader, &info));
        const size_t n = size_t(info.sequences);
        std::vector<uint16_t> blocks(size_t(info.blocks));
        std::vector<int64_t> offset(n), start(n), end(n);
        std::vector<int32_t> size(n);
        check(msa_fasta_copy(loader, blocks.data(), offset.data(), size.data(), start.data(), end.data()));
        database result;
        result.m_entries.reserve(n);
        for (size_t r = 0; r < n; ++r) {
            encoder::buffer b(blocks.begin() + offset[r], blocks.begin() + offset[r] + size[r]);
            std::string description = text.substr(size_t(start[r]), size_t(end[r] - start[r]));
            result.m_indeces[description] = ptrdiff_t(r);
            result.m_entries.push_back({std::move(description), sequence {std::move(b)}});
        }
        return result;
    }

    namespace pairwise
    {
        // ---- distance matrix (src/pairwise/pairwise.cuh:53-117) -----------

        score distance_matrix::operator()(size_t x, size_t y) const
        {
            if (x >= m_count || y >= m_count) throw std::out_of_range("distance matrix offset out of range");
            const size_t hi = std::max(x, y), lo = std::min(x, y);
            return hi > lo ? m_buf[nchoose(hi) + lo] : score {0};
        }

        // ---- scoring tables (src/pairwise/table.cu:254-330) ----------------

        scoring_table::scoring_table(const raw_type& table, element_type penalty)
        :   m_contents {std::make_shared<holder>()}
        ,   m_penalty {penalty}
        {
            for (int i = 0; i < 25; ++i)
                for (int j = 0; j < 25; ++j) m_contents->v[i][j] = table[i][j];
        }

        const std::vector<std::string>& scoring_table::list() noexcept
        {
            static const std::vector<std::string> names = [] {
                std::vector<std::string> v;
                for (int i = 0; i < msa_table_count(); ++i) v.emplace_back(msa_table_name(i));
                return v;
            }();
            return names;
        }

        bool scoring_table::has(const std::string& name)
        {
            const auto& l = list();
            return std::find(l.begin(), l.end(), name) != l.end();
        }

        scoring_table scoring_table::make(const std::string& name)
        {
            int32_t t[625];
            int32_t p = 0;
            if (msa_table_get(name.c_str(), t, &p) != MSA_OK) raise("unknown pairwise scoring table '%s'", name);
            raw_type raw;
            for (int i = 0; i < 625; ++i) raw[i / 25][i % 25] = float(t[i]);
            return scoring_table {raw, float(p)};
        }

        // ---- pairwise::database (src/pairwise/database.cu:24-53) ----------

        database::database(const museqa::database& db)
        {
            for (const auto& e : db) {
                offset.push_back(int64_t(blocks.size()));
                size.push_back(int32_t(e.contents.size()));
                blocks.insert(blocks.end(), e.contents.blocks().begin(), e.contents.blocks().end());
            }
        }

        // ---- algorithms ---------------------------------------------------

        std::vector<pair> algorithm::generate(size_t num) const
        {
            std::vector<pair> pairs;
            pairs.reserve(nchoose(num));
            for (size_t i = 1; i < num; ++i)
                for (size_t j = 0; j < i; ++j) {
                    pair p;
                    p.first = seqref(i);
                    p.second = seqref(j);
                    pairs.push_back(p);
                }
            return pairs;
        }

        namespace
        {
            // The table as int32: the kernels are exact for integer tables,
            // which are all the reference ships.
            void integer_table(const scoring_table& table, int32_t (&out)[625], int32_t& penalty)
            {
                for (int i = 0; i < 625; ++i) {
                    const float v = table.raw()[i / 25][i % 25];
                    if (v != std::nearbyint(v) || std::fabs(v) > 1 << 20)
                        throw exception("scoring table entries must be integers for exact int32 scoring");
                    out[i] = int32_t(v);
                }
                const float p = table.penalty();
                if (p != std::nearbyint(p) || std::fabs(p) > 1 << 20)
                    throw exception("the gap penalty must be an integer for exact int32 scoring");
                penalty = int32_t(p);
            }

            // One kernel family behind the C ABI (hybrid.cu:444-456 is the
            // reference's counterpart).
            struct hip_algorithm : public algorithm
            {
                int kernel;
                explicit hip_algorithm(int k) : kernel {k} {}

                distance_matrix run(const context& ctx) const override
                {
                    const size_t n = ctx.db.count();
                    const database packed {ctx.db};
                    int32_t table[625];
                    int32_t penalty = 0;
                    integer_table(ctx.table, table, penalty);
                    std::vector<int32_t> scores(nchoose(n));
                    msa_nw_options opts {};
                    opts.kernel = kernel;
                    if (!scores.empty()) {
                        int devices = 0;
                        check(msa_device_count(&devices));
                        if (devices < 1) throw exception("no GPU available for the pairwise kernels");
                        check(msa_nw_align(packed.blocks.data(), packed.offset.data(), packed.size.data(),
                                           int32_t(n), nullptr, int64_t(scores.size()), table, penalty, &opts,
                                           scores.data()));
                    }
                    return distance_matrix {std::vector<score>(scores.begin(), scores.end()), n};
                }
            };

            using factory = std::function<std::unique_ptr<algorithm>()>;

            template <int K>
            std::unique_ptr<algorithm> make_hip() { return std::unique_ptr<algorithm>(new hip_algorithm(K)); }

            // src/pairwise/pairwise.cpp:26-35 plus the GPU variants.  Every
            // route gives the scores of sequential.cpp exactly, so the
            // reference's names all take the auto-routed path (the fastest);
            // hip-plain / hip-batched / hip-long force one kernel family.
            const std::vector<std::pair<std::string, factory>>& registry()
            {
                static const std::vector<std::pair<std::string, factory>> r = {
                    {"default", make_hip<MSA_KERNEL_AUTO>},
                    {"needleman", make_hip<MSA_KERNEL_AUTO>},
                    {"hybrid", make_hip<MSA_KERNEL_AUTO>},
                    {"needleman-hybrid", make_hip<MSA_KERNEL_AUTO>},
                    {"sequential", make_hip<MSA_KERNEL_AUTO>},
                    {"distributed", make_hip<MSA_KERNEL_AUTO>},
                    {"needleman-sequential", make_hip<MSA_KERNEL_AUTO>},
                    {"needleman-distributed", make_hip<MSA_KERNEL_AUTO>},
                    {"hip", make_hip<MSA_KERNEL_AUTO>},
                    {"hip-plain", make_hip<MSA_KERNEL_PLAIN>},
                    {"hip-batched", make_hip<MSA_KERNEL_BATCHED>},
                    {"hip-long", make_hip<MSA_KERNEL_LONG>},
                };
                return r;
            }
        }

        bool algorithm::has(const std::string& name)
        {
            for (const auto& e : registry())
                if (e.first == name) return true;
            return false;
        }

        std::unique_ptr<algorithm> algorithm::make(const std::string& name)
        {
            for (const auto& e : registry())
                if (e.first == name) return e.second();
            raise("unknown pairwise algorithm '%s'", name);
        }

        const std::vector<std::string>& algorithm::list() noexcept
        {
            static const std::vector<std::string> names = [] {
                std::vector<std::string> v;
                for (const auto& e : registry()) v.push_back(e.first);
                return v;
            }();
            return names;
        }

        distance_matrix run(const museqa::database& db, const scoring_table& table, const std::string& name)
        {
            const auto worker = algorithm::make(name);
            return worker->run({db, table});
        }
    }

    // ---- phylogeny --------------------------------------------------------

    namespace phylogeny
    {
        const otu& guidetree::operator[](oturef ref) const
        {
            if (ref >= m_nodes.size()) throw exception("guide tree node out of range");
            return m_nodes[ref];
        }

        const otu& guidetree::root() const
        {
            if (m_nodes.empty()) throw exception("the guide tree is empty");
            return m_nodes.back();
        }

        namespace
        {
            static_assert(sizeof(otu) == sizeof(msa_nj_node), "otu mirrors msa_nj_node");

            // Neighbor joining in the HIP kernels behind the C ABI: the
            // reference's sequential semantics (impl/sequential.cpp), either
            // matrix-removal variant.
            struct hip_njoining : public algorithm
            {
                int variant;
                explicit hip_njoining(int v) : variant {v} {}

                guidetree run(const context& ctx) const override
                {
                    const size_t n = ctx.count;
                    if (n > ctx.matrix.count())
                        throw exception("the distance matrix holds fewer sequences than requested");
                    if (n > 65535) throw exception("at most 65535 sequences can be joined");
                    if (n == 0) return guidetree {};
                    std::vector<otu> nodes(2 * n - 1);
                    if (n == 1) {
                        nodes[0] = otu {undefined, {undefined, undefined}, 3.40282347e+38f, 0};
                        return guidetree {std::move(nodes), 1};
                    }
                    // The matrix of the first n sequences is the first n(n-1)/2
                    // entries of the packed buffer (pair order c = i(i-1)/2 + j).
                    int devices = 0;
                    check(msa_device_count(&devices));
                    if (devices < 1) throw exception("no GPU available for neighbor joining");
                    check(msa_nj_build(ctx.matrix.raw().data(), int32_t(n), variant,
                                       reinterpret_cast<msa_nj_node*>(nodes.data())));
                    return guidetree {std::move(nodes), uint32_t(n)};
                }
            };

            using factory = std::function<std::unique_ptr<algorithm>()>;

            template <int V>
            std::unique_ptr<algorithm> make_nj() { return std::unique_ptr<algorithm>(new hip_njoining(V)); }

            // src/phylogeny/phylogeny.cpp:26-41.  The reference's hybrid and
            // sequential variants share one semantics here (the sequential
            // one's pick order, float for float), so each name maps to its
            // matrix-removal variant.
            const std::vector<std::pair<std::string, factory>>& registry()
            {
                static const std::vector<std::pair<std::string, factory>> r = {
                    {"default", make_nj<MSA_NJ_SYMMETRIC>},
                    {"njoining", make_nj<MSA_NJ_SYMMETRIC>},
                    {"hybrid", make_nj<MSA_NJ_SYMMETRIC>},
                    {"linear", make_nj<MSA_NJ_LINEAR>},
                    {"symmetric", make_nj<MSA_NJ_SYMMETRIC>},
                    {"njoining-hybrid", make_nj<MSA_NJ_SYMMETRIC>},
                    {"njoining-linear", make_nj<MSA_NJ_LINEAR>},
                    {"njoining-symmetric", make_nj<MSA_NJ_SYMMETRIC>},
                    {"sequential", make_nj<MSA_NJ_SYMMETRIC>},
                    {"distributed", make_nj<MSA_NJ_SYMMETRIC>},
                    {"njoining-sequential", make_nj<MSA_NJ_SYMMETRIC>},
                    {"njoining-sequential-linear", make_nj<MSA_NJ_LINEAR>},
                    {"njoining-distributed", make_nj<MSA_NJ_SYMMETRIC>},
                    {"njoining-distributed-linear", make_nj<MSA_NJ_LINEAR>},
                    {"hip", make_nj<MSA_NJ_SYMMETRIC>},
                    {"hip-linear", make_nj<MSA_NJ_LINEAR>},
                };
                return r;
            }
        }

        bool algorithm::has(const std::string& name)
        {
            for (const auto& e : registry())
                if (e.first == name) return true;
            return false;
        }

        std::unique_ptr<algorithm> algorithm::make(const std::string& name)
        {
            for (const auto& e : registry())
                if (e.first == name) return e.second();
            raise("unknown phylogeny algorithm '%s'", name);
        }

        const std::vector<std::string>& algorithm::list() noexcept
        {
            static const std::vector<std::string> names = [] {
                std::vector<std::string> v;
                for (const auto& e : registry()) v.push_back(e.first);
                return v;
            }();
            return names;
        }

        guidetree run(const pairwise::distance_matrix& matrix, size_t count, const std::string& name)
        {
            const auto worker = algorithm::make(name);
            return worker->run({matrix, count});
        }
    }

    // ---- pgalign ----------------------------------------------------------

    namespace pgalign
    {
        sequence::sequence(const museqa::sequence& original) : m_original {original}
        {
            const size_t n = original.unpadded();
            m_columns.resize(n + 1);
            for (size_t i = 0; i <= n; ++i) m_columns[i] = uint32_t(i);
        }

        sequence::sequence(const museqa::sequence& original, std::vector<uint32_t> columns, size_t length)
            : m_original {original}, m_columns {std::move(columns)}
        {
            if (m_columns.size() != original.unpadded()) throw exception("one column per residue is needed");
            for (size_t i = 0; i < m_columns.size(); ++i)
                if (m_columns[i] >= length || (i && m_columns[i] <= m_columns[i - 1]))
                    throw exception("residue columns must ascend within the alignment's length");
            m_columns.push_back(uint32_t(length));
        }

        // src/pgalign/sequence.cu:24-50: a binary search over the residues'
        // columns.
        encoder::unit sequence::operator[](ptrdiff_t column) const
        {
            if (column < 0 || size_t(column) >= length()) return encoder::end;
            const auto first = m_columns.begin(), last = m_columns.end() - 1;
            const auto it = std::lower_bound(first, last, uint32_t(column));
            return (it != last && *it == uint32_t(column)) ? m_original[it - first] : encoder::gap;
        }

        std::string sequence::decode() const
        {
            std::string out(length(), encoder::decode(encoder::gap));
            for (size_t i = 0; i + 1 < m_columns.size(); ++i) out[m_columns[i]] = encoder::decode(m_original[i]);
            return out;
        }

        const sequence& alignment::operator[](ptrdiff_t row) const
        {
            if (row < 0 || size_t(row) >= m_rows.size()) throw exception("alignment row out of range");
            return m_rows[row];
        }

        // src/pgalign/alignment.cuh:58-65: the rows [displ, displ + size).
        alignment alignment::slice(ptrdiff_t displ, size_t size) const
        {
            if (displ < 0 || size_t(displ) + size > m_rows.size()) throw exception("alignment slice out of range");
            return alignment {std::vector<sequence>(m_rows.begin() + displ, m_rows.begin() + displ + size)};
        }

        namespace
        {
            // Progressive profile alignment along the guide tree in the HIP
            // kernels behind msa_pg_align (every join optimal under the
            // sum-of-pairs score with the table's linear gap).
            struct hip_progressive : public algorithm
            {
                alignment run(const context& ctx) const override
                {
                    const size_t n = ctx.count;
                    if (n > ctx.db.count()) throw exception("the database holds fewer sequences than requested");
                    if (n == 0) return alignment {};
                    if (ctx.tree.leaves() != n || ctx.tree.count() != 2 * n - 1)
                        throw exception("the guide tree does not join the requested sequences");
                    const pairwise::scoring_table table =
                        ctx.table ? *ctx.table : pairwise::scoring_table::make("default");
                    int32_t t625[625];
                    int32_t penalty = 0;
                    pairwise::integer_table(table, t625, penalty);
                    std::vector<uint16_t> blocks;
                    std::vector<int64_t> offset(n);
                    std::vector<int32_t> size(n);
                    std::vector<size_t> at(n + 1, 0);
                    for (size_t s = 0; s < n; ++s) {
                        const auto& seq = ctx.db[ptrdiff_t(s)].contents;
                        offset[s] = int64_t(blocks.size());
                        size[s] = int32_t(seq.size());
                        blocks.insert(blocks.end(), seq.blocks().begin(), seq.blocks().end());
                        at[s + 1] = at[s] + seq.unpadded();
                    }
                    std::vector<int32_t> columns(std::max<size_t>(at[n], 1));
                    int64_t length = 0;
                    if (n > 1) {
                        int devices = 0;
                        check(msa_device_count(&devices));
                        if (devices < 1) throw exception("no GPU available for the profile aligner");
                    }
                    if (n == 1) {
                        for (size_t i = 0; i < at[1]; ++i) columns[i] = int32_t(i);
                        length = int64_t(at[1]);
                    } else {
                        check(msa_pg_align(blocks.data(), offset.data(), size.data(), int32_t(n),
                                           reinterpret_cast<const msa_nj_node*>(ctx.tree.raw().data()), t625, penalty,
                                           columns.data(), &length, nullptr));
                    }
                    std::vector<sequence> rows;
                    rows.reserve(n);
                    for (size_t s = 0; s < n; ++s)
                        rows.emplace_back(ctx.db[ptrdiff_t(s)].contents,
                                          std::vector<uint32_t>(columns.begin() + at[s], columns.begin() + at[s + 1]),
                                          size_t(length));
                    return alignment {std::move(rows)};
                }
            };

            using factory = std::function<std::unique_ptr<algorithm>()>;

            std::unique_ptr<algorithm> make_progressive() { return std::unique_ptr<algorithm>(new hip_progressive()); }

            // src/pgalign/pgalign.cpp:24-28: every name of the reference runs
            // the GPU progressive aligner (the reference's "myers" is a stub).
            const std::vector<std::pair<std::string, factory>>& registry()
            {
                static const std::vector<std::pair<std::string, factory>> r = {
                    {"default", make_progressive},
                    {"myers", make_progressive},
                    {"sequential", make_progressive},
                    {"myers-sequential", make_progressive},
                    {"hip", make_progressive},
                };
                return r;
            }
        }

        bool algorithm::has(const std::string& name)
        {
            for (const auto& e : registry())
                if (e.first == name) return true;
            return false;
        }

        std::unique_ptr<algorithm> algorithm::make(const std::string& name)
        {
            for (const auto& e : registry())
                if (e.first == name) return e.second();
            raise("unknown pgalign algorithm '%s'", name);
        }

        const std::vector<std::string>& algorithm::list() noexcept
        {
            static const std::vector<std::string> names = [] {
                std::vector<std::string> v;
                for (const auto& e : registry()) v.push_back(e.first);
                return v;
            }();
            return names;
        }

        alignment run(const museqa::database& db, const phylogeny::guidetree& tree, size_t count,
                      const std::string& name)
        {
            const auto worker = algorithm::make(name);
            return worker->run({db, tree, count, nullptr});
        }

        alignment run(const museqa::database& db, const phylogeny::guidetree& tree, size_t count,
                      const pairwise::scoring_table& table, const std::string& name)
        {
            const auto worker = algorithm::make(name);
            return worker->run({db, tree, count, &table});
        }
    }

    // ---- module::pairwise (src/pairwise.cpp:27-55) ------------------------

    namespace
    {
        std::string option(const module::pairwise::options& opts, const std::string& key, const std::string& dflt)
        {
            const auto it = opts.find(key);
            return it == opts.end() ? dflt : it->second;
        }
    }

    bool module::pairwise::check(const options& opts)
    {
        const auto algoname = option(opts, "pairwise", "default");
        if (!museqa::pairwise::algorithm::has(algoname)) raise("unknown pairwise algorithm chosen: '%s'", algoname);
        const auto tablename = option(opts, "scoring-table", "default");
        if (!museqa::pairwise::scoring_table::has(tablename)) raise("unknown scoring table chosen: '%s'", tablename);
        return true;
    }

    module::pairwise::conduit module::pairwise::run(const options& opts, const museqa::database& db)
    {
        const auto table = museqa::pairwise::scoring_table::make(option(opts, "scoring-table", "default"));
        auto result = museqa::pairwise::run(db, table, option(opts, "pairwise", "default"));
        return conduit {db, std::move(result), db.count()};
    }

    // ---- module::phylogeny (src/phylogeny.cpp:25-47) ----------------------

    bool module::phylogeny::check(const options& opts)
    {
        const auto algoname = option(opts, "phylogeny", "default");
        if (!museqa::phylogeny::algorithm::has(algoname)) raise("unknown phylogeny algorithm chosen: '%s'", algoname);
        return true;
    }

    module::phylogeny::conduit module::phylogeny::run(const options& opts, const pairwise::conduit& previous)
    {
        auto tree = museqa::phylogeny::run(previous.distances, previous.count, option(opts, "phylogeny", "default"));
        return conduit {previous.db, std::move(tree)};
    }

    // ---- module::pgalign (src/pgalign.cpp:25-47) --------------------------

    bool module::pgalign::check(const options& opts)
    {
        const auto algoname = option(opts, "pgalign", "default");
        if (!museqa::pgalign::algorithm::has(algoname)) raise("unknown pgalign algorithm chosen: '%s'", algoname);
        const auto tablename = option(opts, "scoring-table", "default");
        if (!museqa::pairwise::scoring_table::has(tablename)) raise("unknown scoring table chosen: '%s'", tablename);
        return true;
    }

    module::pgalign::conduit module::pgalign::run(const options& opts, const phylogeny::conduit& previous)
    {
        const auto table = museqa::pairwise::scoring_table::make(option(opts, "scoring-table", "default"));
        auto aligned = museqa::pgalign::run(previous.db, previous.tree, previous.tree.leaves(), table,
                                            option(opts, "pgalign", "default"));
        return conduit {std::move(aligned)};
    }
}
