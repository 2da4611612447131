// Python binding of the C++ host layer (module msa_amd._museqa).  It plays
// the part of the reference's Cython extensions (src/python/sequence.pyx,
// database.pyx, pairwise.pyx); the Python-facing overloads live in
// msa_amd/museqa/.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "museqa.hpp"

namespace py = pybind11;
using namespace museqa;

PYBIND11_MODULE(_museqa, m)
{
    m.doc() = "C++ host layer of the MI355X pairwise stage (museqa surface)";

    py::register_exception<exception>(m, "MuseqaError", PyExc_RuntimeError);

    m.def("encode_unit", [](const std::string& ch) {
        if (ch.size() != 1) throw std::invalid_argument("expected one character");
        return int(encoder::encode(ch[0]));
    });
    m.def("decode_unit", [](int unit) { return std::string(1, encoder::decode(encoder::unit(unit))); });

    py::class_<sequence>(m, "Sequence")
        .def(py::init<>())
        .def(py::init([](const py::bytes& b) { return sequence {std::string(b)}; }))
        .def("__getitem__", [](const sequence& s, ptrdiff_t i) { return std::string(1, encoder::decode(s[i])); })
        .def("__str__", &sequence::decode)
        .def_property_readonly("length", &sequence::length)
        .def_property_readonly("size", &sequence::size)
        .def("unpadded", &sequence::unpadded)
        .def("blocks", [](const sequence& s) { return s.blocks(); });

    py::class_<database>(m, "Database")
        .def(py::init<>())
        .def("at", [](const database& d, ptrdiff_t i) { const auto& e = d[i]; return py::make_tuple(e.description, e.contents); })
        .def("at", [](const database& d, const std::string& k) { const auto& e = d[k]; return py::make_tuple(e.description, e.contents); })
        .def("add", [](database& d, const std::string& desc, const sequence& s) { d.add(desc, s); })
        .def("add_anonymous", [](database& d, const sequence& s) { d.add(s); })
        .def("merge", &database::merge)
        .def("only_offsets", [](const database& d, const std::vector<ptrdiff_t>& k) {
            return d.only(std::set<ptrdiff_t>(k.begin(), k.end()));
        })
        .def("only_keys", [](const database& d, const std::vector<std::string>& k) {
            return d.only(std::set<std::string>(k.begin(), k.end()));
        })
        .def_static("load", &database::load)
        .def_static("load_on_device", [](const std::string& name) {
            py::gil_scoped_release release;
            return database::load_on_device(name);
        })
        .def_property_readonly("count", &database::count);

    py::class_<pairwise::scoring_table>(m, "ScoringTable")
        .def(py::init([](const std::string& name) { return pairwise::scoring_table::make(name); }))
        .def(py::init([](const std::vector<std::vector<float>>& t, float penalty) {
            if (t.size() != 25) throw std::invalid_argument("a scoring table has 25 rows");
            pairwise::scoring_table::raw_type raw;
            for (int i = 0; i < 25; ++i) {
                if (t[i].size() != 25) throw std::invalid_argument("a scoring table row has 25 entries");
                for (int j = 0; j < 25; ++j) raw[i][j] = t[i][j];
            }
            return pairwise::scoring_table {raw, penalty};
        }))
        .def("at", [](const pairwise::scoring_table& t, size_t x, size_t y) {
            if (x >= 25 || y >= 25) throw exception("scoring table offset out of range");
            return t(x, y);
        })
        .def_property_readonly("penalty", &pairwise::scoring_table::penalty)
        .def_static("list", &pairwise::scoring_table::list)
        .def_static("has", &pairwise::scoring_table::has);

    py::class_<pairwise::distance_matrix>(m, "DistanceMatrix")
        .def(py::init<std::vector<float>, size_t>())
        .def("at", [](const pairwise::distance_matrix& d, size_t x, size_t y) { return d(x, y); })
        .def_property_readonly("count", &pairwise::distance_matrix::count)
        .def("raw", &pairwise::distance_matrix::raw);

    m.def("run", [](const database& db, const pairwise::scoring_table& table, const std::string& algorithm) {
        py::gil_scoped_release release;
        return pairwise::run(db, table, algorithm);
    });
    m.def("algorithms", &pairwise::algorithm::list);
    m.def("has_algorithm", &pairwise::algorithm::has);
    m.def("module_check", &module::pairwise::check);
    m.def("module_run", [](const module::pairwise::options& opts, const database& db) {
        auto c = module::pairwise::run(opts, db);
        return py::make_tuple(c.db, c.distances, c.count);
    });

    // museqa::phylogeny (src/phylogeny/phylogeny.cuh); the reference ships no
    // Python binding for it, this one follows the pairwise binding's shape.
    py::class_<phylogeny::guidetree>(m, "GuideTree")
        .def("node", [](const phylogeny::guidetree& t, phylogeny::oturef ref) {
            const auto& v = t[ref];
            return py::make_tuple(v.parent, py::make_tuple(v.child[0], v.child[1]), v.distance, v.level);
        })
        .def_property_readonly("leaves", &phylogeny::guidetree::leaves)
        .def_property_readonly("count", &phylogeny::guidetree::count)
        .def("root", [](const phylogeny::guidetree& t) {
            t.root();
            return phylogeny::oturef(t.count() - 1);
        })
        .def("raw", [](const phylogeny::guidetree& t) {
            // (parent, child0, child1, distance bits, level) per node, as bytes
            // with the layout of msa_nj_node.
            return py::bytes(reinterpret_cast<const char*>(t.raw().data()), t.raw().size() * sizeof(phylogeny::otu));
        });

    m.def("phylogeny_run", [](const pairwise::distance_matrix& d, size_t count, const std::string& algorithm) {
        py::gil_scoped_release release;
        return phylogeny::run(d, count, algorithm);
    });
    m.def("phylogeny_algorithms", &phylogeny::algorithm::list);
    m.def("phylogeny_has_algorithm", &phylogeny::algorithm::has);
    m.def("phylogeny_module_check", &module::phylogeny::check);
    m.def("phylogeny_module_run", [](const module::phylogeny::options& opts, const database& db,
                                     const pairwise::distance_matrix& d, size_t count) {
        auto c = module::phylogeny::run(opts, module::pairwise::conduit {db, d, count});
        return py::make_tuple(c.db, c.tree);
    });

    // museqa::pgalign (src/pgalign/pgalign.cuh, alignment.cuh, sequence.cuh);
    // the reference ships no Python binding for it either.
    py::class_<pgalign::alignment>(m, "Alignment")
        .def_property_readonly("count", &pgalign::alignment::count)
        .def_property_readonly("length", &pgalign::alignment::length)
        .def("row", [](const pgalign::alignment& a, ptrdiff_t i) {
            const auto& r = a[i];
            const std::vector<uint32_t> cols(r.columns().begin(), r.columns().end() - 1);
            return py::make_tuple(cols, r.decode());
        })
        .def("at", [](const pgalign::alignment& a, ptrdiff_t i, ptrdiff_t column) {
            return std::string(1, encoder::decode(a[i][column]));
        })
        .def("slice", &pgalign::alignment::slice);
    m.def("pgalign_run", [](const database& db, const phylogeny::guidetree& tree, size_t count,
                            const pairwise::scoring_table& table, const std::string& algorithm) {
        py::gil_scoped_release release;
        return pgalign::run(db, tree, count, table, algorithm);
    });
    m.def("pgalign_algorithms", &pgalign::algorithm::list);
    m.def("pgalign_has_algorithm", &pgalign::algorithm::has);
    m.def("pgalign_module_check", &module::pgalign::check);
    m.def("pgalign_module_run", [](const module::pgalign::options& opts, const database& db,
                                   const phylogeny::guidetree& tree) {
        return module::pgalign::run(opts, module::phylogeny::conduit {db, tree}).aligned;
    });
}
