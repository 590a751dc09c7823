// Same-generation Datalog on the parallel-RA program surface (facade over the
// MI355X backend):
//   sg(x,y) :- edge(p,x), edge(p,y), x != y.
//   sg(x,y) :- edge(a,x), sg(a,b), edge(b,y).
// The recursive rule runs as a join chain re-keyed by an acopy, the way Slog
// compiles multi-join bodies.  Usage: same_generation <edge file (tsv_to_bin)>
#include "parallel_RA_inc.h"

static bool not_equal(const u64* const d) { return d[0] != d[1]; }

int main(int argc, char** argv)
{
    mpi_comm mcomm;
    mcomm.create(argc, argv);
    if (argc != 2) {
        std::cout << "usage: same_generation <edge file>" << std::endl;
        MPI_Abort(mcomm.get_comm(), -1);
    }
    relation* edge = new relation(1, true, 2, 300, "edge_2_1_2", argv[1], FULL);
    relation* pairs = new relation(1, true, 2, 301, "inter_body_2_1_2", "", -1);
    relation* sg = new relation(1, true, 2, 302, "sg_2_1_2", "", -1);
    relation* xb = new relation(1, true, 2, 303, "inter_head_2_1_2", "", -1);
    relation* bx = new relation(1, true, 2, 304, "inter_head_2_2_1", "", -1);

    RAM* base = new RAM(false, 1);
    base->add_relation(edge, false);
    base->add_relation(pairs, false);
    base->add_rule(new parallel_join(pairs, edge, FULL, edge, FULL, {1, 3}));

    RAM* filter = new RAM(false, 2);
    filter->add_relation(pairs, true);
    filter->add_relation(sg, false);
    filter->add_rule(new parallel_copy_filter(sg, pairs, DELTA, {0, 1}, not_equal));

    RAM* rec = new RAM(true, 3);
    rec->add_relation(edge, false);
    rec->add_relation(sg, true);
    rec->add_relation(xb, false);
    rec->add_relation(bx, false);
    rec->add_rule(new parallel_join(xb, sg, DELTA, edge, FULL, {3, 1}));
    rec->add_rule(new parallel_acopy(bx, xb, DELTA, {1, 0, 2}));
    rec->add_rule(new parallel_join(sg, bx, DELTA, edge, FULL, {1, 3}));

    LIE* lie = new LIE();
    lie->add_relation(edge);
    lie->add_relation(pairs);
    lie->add_relation(sg);
    lie->add_relation(xb);
    lie->add_relation(bx);
    lie->add_scc(base);
    lie->add_scc(filter);
    lie->add_scc(rec);
    lie->add_scc_dependance(base, filter);
    lie->add_scc_dependance(filter, rec);
    lie->set_comm(mcomm);
    lie->set_batch_size(1);
    lie->execute();
    lie->print_all_relation_size();
    delete lie;
    mcomm.destroy();
    return 0;
}
