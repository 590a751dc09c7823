// The program of tests/transitive_closure.cpp on the edge file argv[1] (its T
// seed empty), with RAM::enable_logging on the fixed-point SCC, against the
// facade: after each batch every relation of that SCC is printed the way
// relation::print prints it.
#include "parallel_RA_inc.h"

int main(int argc, char** argv)
{
    mpi_comm mcomm;
    mcomm.create(argc, argv);
    relation* T = new relation(1, true, 2, 257, "T", argv[2], FULL);
    relation* G = new relation(1, true, 2, 256, "G", argv[1], FULL);
    RAM* fix = new RAM(true, 1);
    fix->add_relation(G, false);
    fix->add_relation(T, true);
    fix->add_rule(new parallel_join(T, T, DELTA, G, FULL, {3, 1}));
    fix->enable_logging();
    RAM* seed = new RAM(false, 2);
    seed->add_relation(G, false);
    seed->add_relation(T, true);
    seed->add_rule(new parallel_copy(T, G, FULL, {1, 0}));
    LIE* lie = new LIE();
    lie->add_relation(T);
    lie->add_relation(G);
    lie->add_scc(fix);
    lie->add_scc(seed);
    lie->add_scc_dependance(seed, fix);
    lie->set_comm(mcomm);
    lie->set_batch_size(1);
    lie->execute();
    lie->print_all_relation_size();
    delete lie;
    mcomm.destroy();
    return 0;
}
