// glp_io.cpp -- the data formats on either side of the phasing run: the
// SNPTools .bin genotype-likelihood reader (GLReader's STBIN path,
// glReader.cpp:25-90,201-285), the position-based transition builder of
// Insti::initialize (insti.cpp:1253-1300) and the VCF writer of
// Insti::save_vcf (insti.cpp:2098-2167).  Host code; gzip files are read and
// written through zlib (plain text is read as well).
#include <zlib.h>

#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/glphase_amd.h"

namespace glp {
int set_error(int code, const char *msg);
}

namespace {

// Reads a gzip (or plain) text file line by line; the '\n' is dropped.
class LineReader {
 public:
  explicit LineReader(const std::string &path) : f_(gzopen(path.c_str(), "rb")) {}
  ~LineReader() {
    if (f_) gzclose(f_);
  }
  bool good() const { return f_ != nullptr; }
  bool getline(std::string &out) {
    out.clear();
    char buf[1 << 16];
    bool any = false;
    while (gzgets(f_, buf, sizeof buf)) {
      any = true;
      size_t n = std::strlen(buf);
      if (n && buf[n - 1] == '\n') {
        out.append(buf, n - 1);
        return true;
      }
      out.append(buf, n);
    }
    return any;
  }

 private:
  gzFile f_;
};

std::vector<std::string> split_tabs(const std::string &s) {
  std::vector<std::string> t;
  size_t b = 0;
  for (;;) {
    const size_t e = s.find('\t', b);
    if (e == std::string::npos) {
      t.push_back(s.substr(b));
      return t;
    }
    t.push_back(s.substr(b, e - b));
    b = e + 1;
  }
}

// Bio::Region("chrom:start-end") (bio.cpp:6-24, bio.hpp:22-68): start and
// end may be empty; 0 is not a valid start.
struct Region {
  std::string chrom;
  unsigned start = 0, end = 0;
  bool empty() const { return chrom.empty(); }
  unsigned startBP() const { return start ? start : 1; }
  unsigned endBP() const { return end ? end : 0xFFFFFFFFu; }
};

Region parse_region(const std::string &s) {
  Region r;
  if (s.empty()) return r;
  const size_t colon = s.rfind(':');
  const size_t dash = colon == std::string::npos ? colon : s.find('-', colon);
  auto digits = [](const std::string &x) {
    for (char c : x)
      if (c < '0' || c > '9') return false;
    return true;
  };
  if (colon == std::string::npos || colon == 0 || dash == std::string::npos)
    throw std::runtime_error("Could not parse region: [" + s + "]");
  r.chrom = s.substr(0, colon);
  const std::string a = s.substr(colon + 1, dash - colon - 1), b = s.substr(dash + 1);
  for (char c : r.chrom)
    if (std::isspace((unsigned char)c)) throw std::runtime_error("Could not parse region: [" + s + "]");
  if (!digits(a) || !digits(b)) throw std::runtime_error("Could not parse region: [" + s + "]");
  if (!a.empty()) r.start = (unsigned)std::stoul(a);
  if (r.start == 0) throw std::range_error("0 is not a valid genomic coordinate");
  if (!b.empty()) r.end = (unsigned)std::stoul(b);
  return r;
}

}  // namespace

struct glp_stbin {
  std::vector<std::string> names;  // kept samples, in file order
  std::vector<std::string> chrom, ref, alt;
  std::vector<unsigned> pos;
  std::vector<float> trios;  // [M][N][3] (RR, Het, AA)
  size_t not_read = 0;
};

namespace {

// GLReader::LoadSTBinNames + LoadFilteredNameIDXs + LoadSTBinGLs.
void read_stbin(const std::string &path, const Region &tr, size_t size_limit,
                const std::string &keep_path, glp_stbin &out) {
  LineReader in(path);
  if (!in.good()) throw std::runtime_error("Error opening 'nameFile' file [" + path + "]");
  std::string line;
  in.getline(line);
  std::vector<std::string> tok = split_tabs(line);
  if (tok.size() < 4)
    throw std::runtime_error("Input bin file [" + path + "] does not contain any sample information");
  const size_t ncol = tok.size() - 3;
  std::unordered_set<std::string> keep;
  if (!keep_path.empty()) {
    LineReader k(keep_path);
    if (!k.good())
      throw std::runtime_error("Error opening 'subset name file' file [" + keep_path + "]");
    std::string n;
    while (k.getline(n)) keep.insert(n);
  }
  std::vector<size_t> idx;
  for (size_t j = 0; j < ncol; ++j)
    if (keep.empty() || keep.count(tok[3 + j])) {
      idx.push_back(j);
      out.names.push_back(tok[3 + j]);
    }
  std::unordered_set<std::string> seen;  // snp_storage's duplicate check (bio.cpp:47-54)
  size_t lineNum = 1;
  while (in.getline(line)) {
    ++lineNum;
    tok = split_tabs(line);
    if (tok.size() != 3 + ncol)
      throw std::runtime_error("Input line " + std::to_string(lineNum) +
                               " does not have the correct number of columns [" +
                               std::to_string(3 + ncol) + "]");
    const unsigned pos = (unsigned)std::stoul(tok[1]);
    if (!tr.empty() && (tr.chrom != tok[0] || tr.startBP() > pos || tr.endBP() < pos)) continue;
    if (size_limit != 0 && out.pos.size() >= size_limit) {
      ++out.not_read;
      continue;
    }
    std::string ref, alt;
    if (tok[2].size() > 2) {
      const size_t sp = tok[2].find_first_of(' ');
      if (sp == std::string::npos) throw std::runtime_error("Could not parse alleles [" + tok[2] + "]");
      ref = tok[2].substr(0, sp);
      alt = tok[2].substr(sp + 1);
    } else {
      ref = std::string(1, tok[2].size() > 0 ? tok[2][0] : '\0');
      alt = std::string(1, tok[2].size() > 1 ? tok[2][1] : '\0');
    }
    // snp_storage_ordered::push_back (bio.cpp:64-75)
    if (!out.pos.empty() && pos < out.pos.back())
      throw std::range_error("Tried to insert position [" + std::to_string(pos) +
                             "] that was less than last position [" +
                             std::to_string(out.pos.back()) + "]");
    if (!out.pos.empty() && out.chrom.front() != tok[0])
      throw std::range_error("Input snp is wrong chromosome [" + tok[0] + "]");
    const std::string key = tok[0] + ":" + std::to_string(pos) + ":" + ref + ":" + alt;
    if (!seen.insert(key).second)
      throw std::runtime_error("Cannot push SNP, it already exists: [" + key + "]");
    out.chrom.push_back(tok[0]);
    out.pos.push_back(pos);
    out.ref.push_back(ref);
    out.alt.push_back(alt);
    // two likelihoods per column: Het then AA (stof, then stof of the rest)
    for (size_t j : idx) {
      const std::string &c = tok[3 + j];
      size_t used = 0;
      const float het = std::stof(c, &used);
      const float aa = std::stof(c.substr(used));
      const float r = 1 - het - aa;
      out.trios.push_back(0.0f < r ? r : 0.0f);  // max(0.0f, 1 - het - aa)
      out.trios.push_back(het);
      out.trios.push_back(aa);
    }
  }
}

// Insti::save_vcf's Phred transform of one probability (insti.cpp:2141-2150).
float phred(float p) {
  if (p == 0) return FLT_MAX_10_EXP * 10;
  if (p == 1) return 0;
  if (p > 1 || p < 0) throw std::runtime_error("probability is not between 0 and 1");
  return -10 * std::log10(p);
}

class Out {
 public:
  explicit Out(const std::string &path) {
    const bool gz = path.size() >= 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    if (gz)
      gz_ = gzopen(path.c_str(), "wb6");
    else
      f_ = std::fopen(path.c_str(), "wb");
    if (!gz_ && !f_) throw std::runtime_error("could not open [" + path + "] for writing");
  }
  ~Out() { close(); }
  void put(const std::string &s) {
    if (gz_) {
      if (!s.empty() && gzwrite(gz_, s.data(), (unsigned)s.size()) != (int)s.size())
        throw std::runtime_error("write failed");
    } else if (std::fwrite(s.data(), 1, s.size(), f_) != s.size()) {
      throw std::runtime_error("write failed");
    }
  }
  void close() {
    if (gz_) gzclose(gz_);
    if (f_) std::fclose(f_);
    gz_ = nullptr;
    f_ = nullptr;
  }

 private:
  gzFile gz_ = nullptr;
  FILE *f_ = nullptr;
};

// A float as an ostream with precision 3 and the default float field prints it.
std::string g3(float v) {
  char b[32];
  std::snprintf(b, sizeof b, "%.3g", (double)v);
  return b;
}

}  // namespace

extern "C" int glp_stbin_read(const char *path, const char *region, size_t size_limit,
                              const char *keep_names_path, glp_stbin **out) {
  if (!path || !out) return glp::set_error(GLP_ERR_ARG, "null path or output");
  *out = nullptr;
  try {
    auto *b = new glp_stbin;
    try {
      read_stbin(path, parse_region(region ? region : ""), size_limit,
                 keep_names_path ? keep_names_path : "", *b);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
    return GLP_OK;
  } catch (const std::exception &e) {
    return glp::set_error(GLP_ERR_ARG, e.what());
  }
}

extern "C" int glp_stbin_dims(const glp_stbin *b, unsigned *num_samples, unsigned *num_sites,
                              size_t *num_not_read) {
  if (!b) return glp::set_error(GLP_ERR_ARG, "null reader");
  if (num_samples) *num_samples = (unsigned)b->names.size();
  if (num_sites) *num_sites = (unsigned)b->pos.size();
  if (num_not_read) *num_not_read = b->not_read;
  return GLP_OK;
}

extern "C" const char *glp_stbin_name(const glp_stbin *b, unsigned i) {
  return b && i < b->names.size() ? b->names[i].c_str() : nullptr;
}

extern "C" int glp_stbin_site(const glp_stbin *b, unsigned m, const char **chrom, unsigned *pos,
                              const char **ref, const char **alt) {
  if (!b || m >= b->pos.size()) return glp::set_error(GLP_ERR_ARG, "site index out of range");
  if (chrom) *chrom = b->chrom[m].c_str();
  if (pos) *pos = b->pos[m];
  if (ref) *ref = b->ref[m].c_str();
  if (alt) *alt = b->alt[m].c_str();
  return GLP_OK;
}

extern "C" int glp_stbin_gls(const glp_stbin *b, int drop_first, float *out) {
  if (!b || !out) return glp::set_error(GLP_ERR_ARG, "null reader or output");
  const size_t n = b->trios.size() / 3;
  for (size_t g = 0; g < n; ++g) {
    const float homR = b->trios[3 * g], het = b->trios[3 * g + 1], homA = b->trios[3 * g + 2];
    if (drop_first) {  // GLReader::drop_first_aware_add_gl_trio (glReader.cpp:201-213)
      const float sum = homR + het + homA;
      out[2 * g] = het / sum;
      out[2 * g + 1] = homA / sum;
    } else {
      out[3 * g] = homR;
      out[3 * g + 1] = het;
      out[3 * g + 2] = homA;
    }
  }
  return GLP_OK;
}

extern "C" void glp_stbin_free(glp_stbin *b) { delete b; }

extern "C" int glp_tran_from_positions(const unsigned *pos, unsigned num_sites, unsigned hn,
                                       double density, double inflation, float *tran_out) {
  if (!pos || !tran_out || num_sites == 0 || hn < 2 || !(density > 0))
    return glp::set_error(GLP_ERR_ARG, "bad argument");
  float mu = 0;
  for (unsigned i = 1; i < hn; i++) mu += 1.0 / i;
  mu = 1 / mu;
  const size_t firstPos = pos[0], lastPos = pos[num_sites - 1];
  if (lastPos < firstPos) return glp::set_error(GLP_ERR_ARG, "positions are not sorted");
  const float rho =
      lastPos == firstPos ? 0.0f : 0.5 * mu * (num_sites - 1) / (lastPos - firstPos) / density;
  tran_out[0] = tran_out[1] = tran_out[2] = 0;
  for (unsigned m = num_sites - 1; m; m--) {
    const float tmp = (pos[m] - pos[m - 1]) * rho;
    float r = tmp / (tmp + hn);
    r *= inflation;
    r = r > 1 ? 1 : r;
    tran_out[m * 3] = (1 - r) * (1 - r);
    tran_out[m * 3 + 1] = r * (1 - r);
    tran_out[m * 3 + 2] = r * r;
  }
  return GLP_OK;
}

extern "C" int glp_write_vcf(const char *path, unsigned num_samples, const char *const *names,
                             unsigned num_sites, const char *const *chrom, const unsigned *pos,
                             const char *const *ref, const char *const *alt,
                             const uint32_t *hsum, unsigned sampling_gens,
                             const char *command_line) {
  if (!path || !names || !chrom || !pos || !ref || !alt || !hsum || num_sites == 0 ||
      sampling_gens == 0)
    return glp::set_error(GLP_ERR_ARG, "bad argument");
  try {
    Out o(path);
    std::string s;
    s += "##fileformat=VCFv4.2\n";
    s += std::string("##source=") + glp_version() + "\n";
    s += std::string("##phaseAndImputeCall=") + (command_line ? command_line : "") + "\n";
    s += std::string("##contig=<ID=") + chrom[0] + ">\n";
    s += "##FORMAT=<ID=GT,Number=1,Type=String,Description=\"Genotype\">\n";
    s += "##FORMAT=<ID=GP,Number=3,Type=Float,Description=\"Phred-scaled genotype posterior "
         "probability\">\n";
    s += "##FORMAT=<ID=APP,Number=2,Type=Float,Description=\"Phred-scaled allelic posterior "
         "probability, P(Allele=1|Haplotype)\">\n";
    s += "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT";
    for (unsigned i = 0; i < num_samples; i++) s += std::string("\t") + names[i];
    o.put(s);
    // Impute::result (impute.cpp:599-613): p = hsum * (1 / sn) in float
    const float norm = 1.0 / sampling_gens;
    for (unsigned m = 0; m < num_sites; m++) {
      s = "\n";
      s += std::string(chrom[m]) + "\t" + std::to_string(pos[m]) + "\t.\t" + ref[m] + "\t" + alt[m] +
           "\t100\tPASS\t.\tGT:GP:APP";
      for (unsigned i = 0; i < num_samples; i++) {
        const float p0 = hsum[(size_t)(2 * i) * num_sites + m] * norm;
        const float p1 = hsum[(size_t)(2 * i + 1) * num_sites + m] * norm;
        const float prr = (1 - p0) * (1 - p1), pra = (1 - p0) * p1 + p0 * (1 - p1),
                    paa = p0 * p1;
        if (prr >= pra && prr >= paa)
          s += "\t0|0";
        else if (pra >= prr && pra >= paa)
          s += p0 > p1 ? "\t1|0" : "\t0|1";
        else
          s += "\t1|1";
        const float v[5] = {phred(prr), phred(pra), phred(paa), phred(p0), phred(p1)};
        for (unsigned k = 0; k < 5; ++k) {
          if (k == 0 || k == 3) s += ":";
          s += g3(v[k]);
          if (k != 2 && k != 4) s += ",";
        }
      }
      o.put(s);
    }
    o.put("\n");
    o.close();
    return GLP_OK;
  } catch (const std::exception &e) {
    return glp::set_error(GLP_ERR_ARG, e.what());
  }
}
