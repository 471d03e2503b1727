// tweakalign -- command-line drop-in for tweakseq's external aligner slot.
//
// tweakseq runs an aligner as a child process (tweakseq/UI/SeqEditMainWin.cpp
// :1616-1663) with an argument list built by AlignmentTool::makeCommand, and
// accepts the result when the process exits 0 and the output file exists
// (:836-861).  This program honours all three argument styles the reference
// builds, so its path can be set in any of the three tool slots:
//   MUSCLE   -in FIN -out FOUT                       (Muscle.cpp:49-53)
//   ClustalO --force -v --outfmt=fa --output-order=tree-order -i FIN -o FOUT
//                                                    (ClustalO.cpp:48-52)
//   MAFFT    --auto --thread -1 FIN   (alignment on stdout; MAFFT.cpp:49-53, :98)
// and the version probes (-version / --version, Muscle.cpp:100-113,
// MAFFT.cpp:103-116, ClustalO.cpp:100-111: the second token is the version).
//
// Input/output are FASTA as tweakseq reads and writes it (FASTAFile.cpp:71-187):
// records start at '>' or ';', the header line is kept verbatim, sequence lines
// are concatenated, output is wrapped at 80 columns.  Gaps ('-', '.') left by
// an earlier alignment are removed first (Sequence::filter keeps them,
// Sequence.cpp:57-69).  Two sequences: Gotoh global alignment on the GPU.
// More: progressive multiple alignment (ta_msa) -- all-vs-all global scores
// on the GPU, a UPGMA guide tree, and profile joins on the GPU.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tweakalign.h"
#include "fasta.h"

namespace {

int fail(const std::string &msg) {
  std::cerr << "tweakalign: " << msg << std::endl;
  return 1;
}

struct Aligner {
  ta_ctx *ctx = nullptr;
  ta_params p{};
  ~Aligner() { if (ctx) ta_destroy(ctx); }
};

}  // namespace

int main(int argc, char **argv) {
  std::string fin, fout;
  bool to_stdout = false, force_dna = false, force_protein = false, fasta_only = false;
  int gap_open = 0, gap_ext = 0;
  for (int k = 1; k < argc; ++k) {
    const std::string a = argv[k];
    auto next = [&](const char *what) -> std::string {
      if (k + 1 >= argc) { fail(std::string("missing value after ") + what); exit(1); }
      return argv[++k];
    };
    if (a == "-version" || a == "--version") {
      std::cout << "tweakalign 1.0.0 (Gotoh on MI355X)" << std::endl;
      return 0;
    } else if (a == "-in" || a == "-i") fin = next(a.c_str());
    else if (a == "-out" || a == "-o") fout = next(a.c_str());
    else if (a == "-dna") force_dna = true;
    else if (a == "-protein") force_protein = true;
    else if (a == "-gapopen") gap_open = atoi(next("-gapopen").c_str());
    else if (a == "-gapextend") gap_ext = atoi(next("-gapextend").c_str());
    else if (a == "--fasta-roundtrip") fasta_only = true;  // parse and rewrite only (no GPU)
    else if (a == "--thread") next("--thread");             // MAFFT style, ignored
    else if (a == "--auto" || a == "--force" || a == "-v" || a.rfind("--outfmt", 0) == 0 ||
             a.rfind("--output-order", 0) == 0) {
    }  // ClustalO / MAFFT flags with no meaning here
    else if (!a.empty() && a[0] != '-' && fin.empty()) { fin = a; to_stdout = true; }  // MAFFT: input only
    else return fail("unknown argument " + a);
  }
  if (fin.empty()) return fail("no input file (-in FILE)");
  if (fout.empty()) to_stdout = true;

  std::ifstream in(fin);
  if (!in) return fail("cannot open " + fin);
  std::vector<Record> recs;
  read_fasta(in, recs);
  if (recs.empty()) return fail("no sequences in " + fin);
  for (Record &r : recs) r.seq = dealign(r.seq);

  std::vector<Record> outrecs = recs;
  if (!fasta_only && recs.size() >= 2) {
    Aligner al;
    const bool dna = force_dna || (!force_protein && looks_dna(recs));
    al.p.mode = TA_GLOBAL;
    al.p.alphabet = dna ? TA_DNA : TA_PROTEIN;
    al.p.gap_open = gap_open ? gap_open : (dna ? -5 : -10);
    al.p.gap_extend = gap_ext ? gap_ext : (dna ? -2 : -1);
    al.p.match = 2;
    al.p.mismatch = -3;
    al.p.lanes = TA_LANES_AUTO;
    int rc = ta_create(&al.ctx, 0);
    if (rc) return fail(std::string("GPU: ") + ta_strerror(rc));
    const size_t N = recs.size();
    if (N == 2) {
      // Two rows (the editor's "realign selection"): the single-pair path.
      // A CIGAR run of length L consumes L residues in at most L + 1 <= 2L
      // characters, so 2 (la + lb) bytes always suffice: one call.
      const std::string &sa = recs[0].seq, &sb = recs[1].seq;
      ta_result res{};
      std::vector<char> cig(2 * (sa.size() + sb.size()) + 1);
      int64_t used = 0;
      al.p.traceback = 1;
      rc = ta_align_pair(al.ctx, &al.p, sa.data(), (int32_t)sa.size(), sb.data(), (int32_t)sb.size(), &res,
                         cig.data(), (int64_t)cig.size(), &used);
      if (rc) return fail(std::string("alignment failed: ") + ta_strerror(rc));
      const std::string ops = expand(cig.data() + res.cigar_off, res.cigar_len);
      // A global alignment's ops consume both sequences exactly; anything
      // else is an engine failure, reported as a failed run (exit status 1,
      // no output file) rather than written as a misaligned FASTA.
      if (!consumes(ops, recs[0].seq.size(), recs[1].seq.size()))
        return fail("alignment failed: the CIGAR does not consume both sequences");
      std::string ra, rb;
      render(recs[0].seq, recs[1].seq, ops, ra, rb);
      outrecs[0].seq = ra;
      outrecs[1].seq = rb;
    } else {
      std::vector<const char *> seqs(N);
      std::vector<int32_t> lens(N);
      size_t total = 0;
      for (size_t x = 0; x < N; ++x) {
        seqs[x] = recs[x].seq.data();
        lens[x] = (int32_t)recs[x].seq.size();
        total += recs[x].seq.size();
      }
      std::vector<int32_t> cols(std::max<size_t>(total, 1));
      int32_t alen = 0;
      al.p.traceback = 0;
      rc = ta_msa(al.ctx, &al.p, seqs.data(), lens.data(), (int32_t)N, &alen, cols.data(), nullptr);
      if (rc) return fail(std::string("alignment failed: ") + ta_strerror(rc));
      size_t k = 0;
      for (size_t x = 0; x < N; ++x) {
        std::string row((size_t)alen, '-');
        for (char ch : recs[x].seq) row[(size_t)cols[k++]] = ch;
        outrecs[x].seq = row;
      }
    }
  }

  if (to_stdout) {
    write_fasta(std::cout, outrecs);
    std::cout.flush();
    return std::cout ? 0 : 1;
  }
  std::ofstream out(fout);
  if (!out) return fail("cannot write " + fout);
  write_fasta(out, outrecs);
  out.close();
  return out ? 0 : 1;
}
