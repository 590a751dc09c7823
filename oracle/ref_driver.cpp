// Test infrastructure: a program driver written against the reference's own
// C++ API (relation, RAM, parallel_join / parallel_copy / parallel_acopy /
// parallel_copy_filter, LIE), compiled together with the reference's sources
// by `make -C oracle refbuild` into oracle/_ref/ref_driver, and run under
// mpirun on the CPU to write golden vectors for the oracle.  Nothing of the
// product links or runs it.
//
// Usage: mpirun -n <ranks> ref_driver <spec> <output dir>
//
// The spec is plain text, one directive per line (relations, SCCs and rules
// are numbered in the order they appear):
//   rel <name> <arity> <jcc> <tag> <tsv_to_bin file>
//   scc <id> <fixpoint 0|1>
//   sccrel <scc#> <rel#> <move_full_to_delta 0|1>
//   join <scc#> <out#> <in0#> <ver0> <in1#> <ver1> <projection...>
//   copy <scc#> <out#> <in#> <ver> <reorder...>
//   acopy <scc#> <out#> <in#> <ver> <reorder...>
//   filter <scc#> <out#> <in#> <ver> <op> <reorder...>   op: 0 EQ 1 NEQ 2 LT 3 LE 4 GT 5 GE on columns 0, 1
//   dep <scc#> <scc#>
//   batch <batch size>
//   log <scc#>                 RAM::enable_logging: every relation of the SCC printed after each batch
// Versions are 0 = DELTA, 1 = FULL (src/relation/balanced_hash_relation.h:13).
//
// Output: the reference's own checkpoint dumps, one per batch (LIE::execute
// with cp_iteration 1 and separate per-rank files, src/lie/lie.cpp:353-423),
// under <output dir>/checkpoint-<scc id>-<loop>/, and after execute() every
// relation's full version written by relation::parallel_IO into
// <output dir>/final/ (<name>_full_<rank> rows in the rank's trie order).
#include "parallel_RA_inc.h"

#include <fstream>
#include <sstream>
#include <sys/stat.h>

namespace {

bool f_eq(const u64* const d) { return d[0] == d[1]; }
bool f_neq(const u64* const d) { return d[0] != d[1]; }
bool f_lt(const u64* const d) { return d[0] < d[1]; }
bool f_le(const u64* const d) { return d[0] <= d[1]; }
bool f_gt(const u64* const d) { return d[0] > d[1]; }
bool f_ge(const u64* const d) { return d[0] >= d[1]; }
bool (*const kFilters[6])(const u64* const) = {f_eq, f_neq, f_lt, f_le, f_gt, f_ge};

std::vector<int> rest_ints(std::istringstream& in)
{
    std::vector<int> v;
    int x;
    while (in >> x) v.push_back(x);
    return v;
}

} // namespace

int main(int argc, char** argv)
{
    mpi_comm mcomm;
    mcomm.create(argc, argv);
    if (argc != 3) {
        if (mcomm.get_rank() == 0) std::cerr << "usage: ref_driver <spec> <output dir>\n";
        MPI_Abort(mcomm.get_comm(), 2);
    }
    std::ifstream spec(argv[1]);
    if (!spec) {
        std::cerr << "cannot open " << argv[1] << "\n";
        MPI_Abort(mcomm.get_comm(), 2);
    }
    const std::string out = argv[2];

    std::vector<relation*> rels;
    std::vector<RAM*> sccs;
    std::vector<std::pair<int, int>> deps;
    u32 batch = 1;
    std::string line;
    while (std::getline(spec, line)) {
        std::istringstream in(line);
        std::string op;
        if (!(in >> op) || op[0] == '#') continue;
        if (op == "rel") {
            std::string name, file;
            u32 arity, jcc, tag;
            in >> name >> arity >> jcc >> tag >> file;
            rels.push_back(new relation(jcc, true, arity, tag, name, file, FULL));
        } else if (op == "scc") {
            int id, fix;
            in >> id >> fix;
            sccs.push_back(new RAM(fix != 0, id));
        } else if (op == "sccrel") {
            int s, r, st;
            in >> s >> r >> st;
            sccs[s]->add_relation(rels[r], st != 0);
        } else if (op == "join") {
            int s, o, a, va, b, vb;
            in >> s >> o >> a >> va >> b >> vb;
            sccs[s]->add_rule(new parallel_join(rels[o], rels[a], va, rels[b], vb, rest_ints(in)));
        } else if (op == "copy" || op == "acopy") {
            int s, o, a, va;
            in >> s >> o >> a >> va;
            if (op == "copy")
                sccs[s]->add_rule(new parallel_copy(rels[o], rels[a], va, rest_ints(in)));
            else
                sccs[s]->add_rule(new parallel_acopy(rels[o], rels[a], va, rest_ints(in)));
        } else if (op == "filter") {
            int s, o, a, va, f;
            in >> s >> o >> a >> va >> f;
            sccs[s]->add_rule(new parallel_copy_filter(rels[o], rels[a], va, rest_ints(in), kFilters[f]));
        } else if (op == "dep") {
            int a, b;
            in >> a >> b;
            deps.emplace_back(a, b);
        } else if (op == "batch") {
            in >> batch;
        } else if (op == "log") {
            int a;
            in >> a;
            sccs[a]->enable_logging();
        } else {
            std::cerr << "unknown directive " << op << "\n";
            MPI_Abort(mcomm.get_comm(), 2);
        }
    }

    LIE* lie = new LIE();
    for (relation* r : rels) lie->add_relation(r);
    for (RAM* s : sccs) lie->add_scc(s);
    for (auto& d : deps) lie->add_scc_dependance(sccs[d.first], sccs[d.second]);
    lie->enable_separate_io();
    lie->set_cp_iteration(1);
    lie->set_output_dir(out);
    lie->enable_IO();
    lie->set_name(argv[0]);
    lie->set_comm(mcomm);
    lie->set_batch_size(batch);
    lie->execute();
    lie->print_all_relation_size();

    const std::string fin = out + "/final";
    if (mcomm.get_rank() == 0) mkdir(fin.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    MPI_Barrier(mcomm.get_comm());
    for (relation* r : rels) {
        r->set_separate_io(true);
        r->parallel_IO(fin);
    }
    delete lie;
    mcomm.destroy();
    return 0;
}
