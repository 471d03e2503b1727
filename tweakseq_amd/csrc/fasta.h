// FASTA records and CIGAR helpers of the tweakalign CLI (cli.cpp), kept in
// a header of their own so the host-side sanitizer build of the tests can
// exercise them without a GPU.
//
// FASTA as tweakseq reads and writes it (tweakseq/Core/FASTAFile.cpp:71-187):
// records start at '>' or ';', the header line is kept verbatim, sequence
// lines are concatenated, output is wrapped at 80 columns.
#pragma once
#include <cctype>
#include <istream>
#include <ostream>
#include <string>
#include <vector>

namespace {

struct Record {
  std::string header;  // verbatim, including '>' or ';'
  std::string seq;
};

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && isspace((unsigned char)s[a])) ++a;
  while (b > a && isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

// FASTAFile::read rules (FASTAFile.cpp:71-147).
bool read_fasta(std::istream &in, std::vector<Record> &out) {
  enum { SEEK, COMMENT, SEQ } st = SEEK;
  std::string line;
  while (std::getline(in, line)) {
    const std::string s = trim(line);
    if (s.empty()) continue;
    const char c = s[0];
    if (st == SEEK) {
      if (c == ';' || c == '>') { out.push_back({s, ""}); st = COMMENT; }
    } else if (st == COMMENT) {
      if (c == ';') continue;
      out.back().seq = s;
      st = SEQ;
    } else {
      if (c == ';' || c == '>') { out.push_back({s, ""}); st = COMMENT; }
      else out.back().seq += s;
    }
  }
  return true;
}

// FASTAFile::write rules (FASTAFile.cpp:149-171).
void write_fasta(std::ostream &o, const std::vector<Record> &recs) {
  for (const Record &r : recs) {
    o << r.header << '\n';
    for (size_t p = 0; p < r.seq.size(); p += 80) o << r.seq.substr(p, 80) << '\n';
  }
}

std::string dealign(const std::string &s) {
  std::string r;
  for (char c : s)
    if (c != '-' && c != '.') r.push_back(c);
  return r;
}

bool looks_dna(const std::vector<Record> &recs) {
  size_t n = 0;
  for (const Record &r : recs)
    for (char c : r.seq) {
      const char u = (char)toupper((unsigned char)c);
      if (u != 'A' && u != 'C' && u != 'G' && u != 'T') return false;
      ++n;
    }
  return n > 0;
}

// CIGAR text -> ops string (M / I / D per column).
std::string expand(const char *cig, int len) {
  std::string ops;
  int v = 0;
  for (int p = 0; p < len; ++p) {
    if (isdigit((unsigned char)cig[p])) { v = v * 10 + (cig[p] - '0'); continue; }
    ops.append((size_t)v, cig[p]);
    v = 0;
  }
  return ops;
}

// Gapped rows of a global alignment of a (query) and b (reference).
void render(const std::string &a, const std::string &b, const std::string &ops, std::string &ra,
            std::string &rb) {
  size_t i = 0, j = 0;
  ra.clear();
  rb.clear();
  for (char o : ops) {
    if (o == 'M') { ra += a[i++]; rb += b[j++]; }
    else if (o == 'I') { ra += a[i++]; rb += '-'; }
    else { ra += '-'; rb += b[j++]; }
  }
}

// True when the ops of a global alignment consume exactly la query ('M',
// 'I') and lb reference ('M', 'D') residues: a CIGAR that does not is an
// engine failure, never an alignment to write.
bool consumes(const std::string &ops, size_t la, size_t lb) {
  size_t i = 0, j = 0;
  for (char o : ops) {
    if (o != 'M' && o != 'I' && o != 'D') return false;
    i += o != 'D';
    j += o != 'I';
  }
  return i == la && j == lb;
}

}  // namespace
