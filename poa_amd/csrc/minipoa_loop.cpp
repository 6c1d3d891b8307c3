// minipoa's main loop (src/main.cpp:84-126) written against the C++ facade
// include/minipoa_amd.hpp: the program a minipoa user keeps, with the alignment path on the
// GPU. Reads one FASTA/FASTQ(.gz) file and, with -i, a plain GFA to align it to; flags -M -X -O -E
// -b -f -B -S -k -w -p -r -i as minipoa's (src/main.cpp:14-35). Prints what minipoa prints.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "minipoa_amd.hpp"

using namespace minipoa_amd;

int main(int argc, char** argv) {
  para_t* para = new para_t();
  const char* path = nullptr;
  const char* inc = nullptr;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto val = [&]() { return i + 1 < argc ? std::atoi(argv[++i]) : 0; };
    if (a == "-M") para->match = val();
    else if (a == "-X") para->mismatch = val();
    else if (a == "-O") para->gap_open1 = val();
    else if (a == "-E") para->gap_ext1 = val();
    else if (a == "-b") para->b = val();
    else if (a == "-f") para->f = val();
    else if (a == "-B") para->ab_band = true;
    else if (a == "-S") para->enable_seeding = true;
    else if (a == "-p") para->progressive_poa = true;
    else if (a == "-k") para->k = val();
    else if (a == "-w") para->mm_w = val();
    else if (a == "-r") para->result = val();
    else if (a == "-i") inc = i + 1 < argc ? argv[++i] : nullptr;
    else path = argv[i];
  }
  if (!path) {
    std::cerr << "usage: minipoa_loop [flags] [-i graph.gfa] reads.fa\n";
    return 1;
  }
  try {
    graph* DAG = new graph();
    DAG->init(para);
    std::vector<seq_t> seqs;
    if (inc) seqs = read_gfa(para, DAG, inc);
    const int exist_seq_num = (int)seqs.size();
    readFile(seqs, path);
    minimizer_t* mm = new minimizer_t(para, seqs);
    if (!inc) mm->get_guide_tree(para);
    for (int i = exist_seq_num; i < (int)seqs.size(); i++) {
      const int rid = mm->ord[i];
      std::vector<res_t> res =
          alignment(para, DAG, mm, rid, seqs[rid].seq.c_str(), (int)seqs[rid].seq.size());
      DAG->add_path(para->m, i, res, 1);
      DAG->topsort(para, i + 1 == (int)seqs.size());
    }
    if (para->result == 0) DAG->output_consensus();
    else if (para->result == 1) DAG->output_rc_msa(para, mm->rid_to_ord, seqs);
    else if (para->result == 2) DAG->output_gfa(mm->rid_to_ord, seqs);
    else if (para->result == 3) std::cout << DAG->text(3, &seqs, &mm->rid_to_ord);
    delete mm;
    delete DAG;
  } catch (const error& e) {
    // minipoa exits 0 on a band violation (align.cpp:1584-1588) and 1 on a backtrack error
    std::cerr << e.what() << "\n";
    return e.code == MPOA_E_BAND ? 0 : 1;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
  delete para;
  return 0;
}
