// C++ host layer of the MI355X pairwise stage: the surface of the reference's
// museqa::pairwise path, re-implemented over the C ABI of libmsa_hip.so.
//
// What it keeps from the reference (names, argument meaning, error
// behaviour), each restated in code of this project:
//   encoder        src/encoder.hpp:18-83, src/encoder.cpp:27-129
//   sequence       src/sequence.hpp:24-188
//   database       src/database.hpp:25-208, src/database.cpp:24-53
//   FASTA loading  src/io/loader/parser/fasta.cpp:26-70
//   pairwise::{score, seqref, pair, distance_matrix, scoring_table, context,
//              algorithm, run}   src/pairwise/pairwise.cuh:24-244,
//                                src/pairwise/pairwise.cpp:26-85,
//                                src/pairwise/table.cu:254-330
//   pairwise::database           src/pairwise/database.cuh:27-91, database.cu:24-72
//   module::pairwise::{run, check}  src/pairwise.cpp:27-55
//   phylogeny::{otu, guidetree, context, algorithm, run}
//                                src/phylogeny/phylogeny.cuh:24-134,
//                                src/phylogeny/phylogeny.cpp:26-75,
//                                src/phylogeny/tree.cuh
//   module::phylogeny::{run, check} src/phylogeny.cpp:25-47
// The algorithm registries hold the reference's names plus the GPU variants;
// every name runs on the GPU (algorithm::list() gives them).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace museqa
{
    // museqa::exception (src/exception.hpp): a runtime error with a message.
    struct exception : public std::runtime_error
    {
        using std::runtime_error::runtime_error;
    };

    namespace encoder
    {
        using unit = uint8_t;
        using block = uint16_t;
        using buffer = std::vector<block>;

        enum : unit { end = 0x18 };
        enum : unit { gap = 0x19 };
        enum : size_t { block_size = 3 };

        unit access(block tgt, uint8_t offset) noexcept;
        unit access(const block* tgt, size_t offset) noexcept;
        unit encode(char) noexcept;
        buffer encode(const char*, size_t);
        char decode(unit);
        std::string decode(const block*, size_t nblocks);
    }

    class sequence
    {
        public:
            static constexpr encoder::unit padding = encoder::end;

            sequence() = default;
            sequence(const encoder::buffer& blocks) : m_blocks {blocks} {}
            sequence(encoder::buffer&& blocks) noexcept : m_blocks {std::move(blocks)} {}
            sequence(const char* ptr, size_t size) : m_blocks {encoder::encode(ptr, size)} {}
            sequence(const std::string& str) : sequence {str.data(), str.size()} {}

            // Code at a residue offset; throws std::out_of_range past the padded length.
            encoder::unit operator[](ptrdiff_t offset) const;
            encoder::block block(ptrdiff_t offset) const { return m_blocks.at(offset); }
            size_t size() const noexcept { return m_blocks.size(); }          // blocks
            size_t length() const noexcept { return m_blocks.size() * encoder::block_size; }
            size_t unpadded() const noexcept;
            std::string decode() const { return encoder::decode(m_blocks.data(), m_blocks.size()); }
            const encoder::buffer& blocks() const noexcept { return m_blocks; }

        private:
            encoder::buffer m_blocks;
    };

    class database;
    namespace io { namespace parser { database fasta_on_device(const std::string& text); } }

    class database
    {
        public:
            struct entry_type
            {
                std::string description;
                sequence contents;
            };

            database() = default;

            const entry_type& operator[](ptrdiff_t offset) const;
            const entry_type& operator[](const std::string& key) const;
            void add(const std::string& description, const sequence& elem);
            void add(const sequence& elem);
            void add(const std::vector<sequence>& elems);
            database only(const std::set<ptrdiff_t>& keys) const;
            database only(const std::set<std::string>& keys) const;
            void merge(const database& db);
            size_t count() const noexcept { return m_entries.size(); }
            auto begin() const noexcept { return m_entries.begin(); }
            auto end() const noexcept { return m_entries.end(); }

            // io::loader<database> for ".fa" / ".fasta" files (src/io/loader/database.cpp).
            static database load(const std::string& filename);
            // The same database, parsed on the GPU (msa_fasta_*): identical
            // records and contents; ".fa" / ".fasta" files only.
            static database load_on_device(const std::string& filename);

        private:
            friend database io::parser::fasta_on_device(const std::string&);
            static int anonymous;
            std::vector<entry_type> m_entries;
            std::map<std::string, ptrdiff_t> m_indeces;
    };

    namespace io
    {
        namespace parser
        {
            database fasta(const std::string& filename);
            // FASTA text already in memory, parsed on the GPU.
            database fasta_on_device(const std::string& text);
        }
    }

    namespace pairwise
    {
        using score = float;
        using seqref = int_least32_t;

        union pair
        {
            struct { seqref first, second; };
            seqref id[2];
        };

        class distance_matrix
        {
            public:
                using element_type = score;

                distance_matrix() = default;
                distance_matrix(std::vector<score> buf, size_t count) : m_buf {std::move(buf)}, m_count {count} {}

                // [x, y] reads buf[nchoose(max) + min]; the diagonal is 0.
                score operator()(size_t x, size_t y) const;
                size_t count() const noexcept { return m_count; }
                const std::vector<score>& raw() const noexcept { return m_buf; }

            private:
                std::vector<score> m_buf;
                size_t m_count = 0;
        };

        class scoring_table
        {
            public:
                using element_type = score;
                using raw_type = element_type[25][25];

                scoring_table() = default;
                scoring_table(const raw_type& table, element_type penalty);

                element_type operator()(size_t x, size_t y) const { return m_contents->v[x][y]; }
                element_type penalty() const noexcept { return m_penalty; }
                const raw_type& raw() const noexcept { return m_contents->v; }

                static bool has(const std::string&);
                static scoring_table make(const std::string&);
                static const std::vector<std::string>& list() noexcept;

            private:
                struct holder { raw_type v; };
                std::shared_ptr<holder> m_contents;
                element_type m_penalty = 0;
        };

        // All sequences back to back plus per-sequence (offset, size) in blocks.
        struct database
        {
            std::vector<uint16_t> blocks;
            std::vector<int64_t> offset;
            std::vector<int32_t> size;

            database() = default;
            explicit database(const museqa::database& db);
            size_t count() const noexcept { return size.size(); }
        };

        struct context
        {
            const museqa::database& db;
            const scoring_table& table;
        };

        struct algorithm
        {
            virtual ~algorithm() = default;
            virtual std::vector<pair> generate(size_t) const;
            virtual distance_matrix run(const context&) const = 0;

            static bool has(const std::string&);
            static std::unique_ptr<algorithm> make(const std::string&);
            static const std::vector<std::string>& list() noexcept;
        };

        distance_matrix run(const museqa::database& db, const scoring_table& table,
                            const std::string& algorithm = "default");

        inline size_t nchoose(size_t n) noexcept { return n ? n * (n - 1) / 2 : 0; }
    }

    namespace phylogeny
    {
        using oturef = uint32_t;

        // One node of the guide tree (src/phylogeny/tree.cuh, star.cuh:63-80):
        // leaves are 0..count-1, joins are numbered count, count+1, ... in
        // join order and the root is the last node.
        struct otu
        {
            oturef parent;
            oturef child[2];                // ordered by id; undefined for a leaf
            float distance;                 // branch length to the parent; FLT_MAX for the root
            uint32_t level;                 // 0 for a leaf, 1 + max(children) above
        };

        enum : oturef { undefined = 0xFFFFFFFFu };

        class guidetree
        {
            public:
                guidetree() = default;
                guidetree(std::vector<otu> nodes, uint32_t leaves) : m_nodes {std::move(nodes)}, m_leaves {leaves} {}

                const otu& operator[](oturef ref) const;
                const otu& root() const;
                uint32_t leaves() const noexcept { return m_leaves; }
                size_t count() const noexcept { return m_nodes.size(); }
                const std::vector<otu>& raw() const noexcept { return m_nodes; }

            private:
                std::vector<otu> m_nodes;
                uint32_t m_leaves = 0;
        };

        // src/phylogeny/phylogeny.cuh:77-112
        struct context
        {
            const pairwise::distance_matrix& matrix;
            const size_t count;
        };

        struct algorithm
        {
            virtual ~algorithm() = default;
            virtual guidetree run(const context&) const = 0;

            static bool has(const std::string&);
            static std::unique_ptr<algorithm> make(const std::string&);
            static const std::vector<std::string>& list() noexcept;
        };

        // src/phylogeny/phylogeny.cuh:114-134
        guidetree run(const pairwise::distance_matrix& matrix, size_t count,
                      const std::string& algorithm = "default");
    }

    namespace pgalign
    {
        // A sequence with gaps (src/pgalign/sequence.cuh): the sequence's
        // first unpadded() residues and, per residue, its column, ascending;
        // the last entry of columns() is the row's length, gaps included.
        class sequence
        {
            public:
                sequence() = default;
                // The sequence alone: residue i in column i.
                explicit sequence(const museqa::sequence& original);
                // The sequence placed in an alignment: one column per residue
                // and the alignment's length.
                sequence(const museqa::sequence& original, std::vector<uint32_t> columns, size_t length);

                // Code at a column: the residue's code, encoder::gap between
                // residues, encoder::end past the row's length
                // (src/pgalign/sequence.cu:24-50).
                encoder::unit operator[](ptrdiff_t column) const;
                size_t length() const noexcept { return m_columns.empty() ? 0 : m_columns.back(); }
                size_t residues() const noexcept { return m_columns.empty() ? 0 : m_columns.size() - 1; }
                std::string decode() const;
                const std::vector<uint32_t>& columns() const noexcept { return m_columns; }
                const museqa::sequence& original() const noexcept { return m_original; }

            private:
                museqa::sequence m_original;
                std::vector<uint32_t> m_columns;
        };

        // The alignment of one or more sequences (src/pgalign/alignment.cuh):
        // one gapped row per sequence, in database order.
        class alignment
        {
            public:
                alignment() = default;
                explicit alignment(std::vector<sequence> rows) : m_rows {std::move(rows)} {}

                const sequence& operator[](ptrdiff_t row) const;
                size_t count() const noexcept { return m_rows.size(); }
                size_t length() const noexcept { return m_rows.empty() ? 0 : m_rows.front().length(); }
                alignment slice(ptrdiff_t displ, size_t size) const;

            private:
                std::vector<sequence> m_rows;
        };

        // src/pgalign/pgalign.cuh:22-31, plus the scoring table the joins
        // use (the reference's stub needs none; null = the "default" table).
        struct context
        {
            const museqa::database& db;
            const phylogeny::guidetree& tree;
            const size_t count;
            const pairwise::scoring_table* table = nullptr;
        };

        struct algorithm
        {
            virtual ~algorithm() = default;
            virtual alignment run(const context&) const = 0;

            static bool has(const std::string&);
            static std::unique_ptr<algorithm> make(const std::string&);
            static const std::vector<std::string>& list() noexcept;
        };

        // src/pgalign/pgalign.cuh:82-90.
        alignment run(const museqa::database& db, const phylogeny::guidetree& tree, size_t count,
                      const std::string& algorithm = "default");
        alignment run(const museqa::database& db, const phylogeny::guidetree& tree, size_t count,
                      const pairwise::scoring_table& table, const std::string& algorithm = "default");
    }

    namespace module
    {
        // The pipeline module contract: option values in, conduit out.
        struct pairwise
        {
            struct conduit
            {
                museqa::database db;
                museqa::pairwise::distance_matrix distances;
                size_t count;
            };

            using options = std::map<std::string, std::string>;
            static const char* name() { return "pairwise"; }
            static bool check(const options& opts);
            static conduit run(const options& opts, const museqa::database& db);
        };

        // src/phylogeny.cpp:25-47: the pairwise conduit in, the tree out.
        struct phylogeny
        {
            struct conduit
            {
                museqa::database db;
                museqa::phylogeny::guidetree tree;
            };

            using options = std::map<std::string, std::string>;
            static const char* name() { return "phylogeny"; }
            static bool check(const options& opts);
            static conduit run(const options& opts, const pairwise::conduit& previous);
        };

        // src/pgalign.cpp:25-47: the phylogeny conduit in, the alignment out;
        // the joins score with the option "scoring-table" (as the pairwise
        // module reads it).
        struct pgalign
        {
            struct conduit
            {
                museqa::pgalign::alignment aligned;
            };

            using options = std::map<std::string, std::string>;
            static const char* name() { return "pgalign"; }
            static bool check(const options& opts);
            static conduit run(const options& opts, const phylogeny::conduit& previous);
        };
    }
}
