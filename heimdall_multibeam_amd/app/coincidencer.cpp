// coincidencer -- multibeam coincidence of the search's candidates, for the
// candidate lines heimdall writes to .cand files or sends over TCP
// (Applications/coincidencer.C, Applications/Candidates.C).
//
//   coincidencer [-n nbeams] [-t sep_time] [-f sep_filter] [-v] beam_files...
//       reads the .cand files of one gulp (one per beam) and writes
//       <first_sample_utc>_all.cand, the prefix of the first file's name up
//       to its first '_'.
//   coincidencer -p port [-a address] [-n nbeams] [-t sep_time] [-f sep_filter] [-v]
//       listens for heimdall's TCP messages (one connection per gulp of a
//       beam group: a header "utc_start first_sample_utc first_sample first_beam
//       nbeams n_events" then n_events candidate lines), gathers the beams of
//       each gulp and writes <utc_start>/<first_sample_utc>_all.cand once all
//       nbeams beams arrived (or more than 4 gulps are waiting).
//
// Coincidence (Candidates.C:63-77, 257-301): candidate j of beam i counts one
// more beam for every other beam k holding a candidate l with
//   |sample_idx_l - sample_idx_j| <= int64(sep_time * 2^max(filter_l, filter_j)),
//   |filter_l - filter_j| <= sep_filter, |dm_trial_l - dm_trial_j| <= 9999 and
//   |snr_l - snr_j| / (snr_l + snr_j) <= 0.4;
// only the first such l of each beam k counts, and it becomes the primary
// beam when its S/N beats the best so far.  Output lines (Candidates.C:80-87):
//   snr sample_idx time filter dm_trial dm members begin end nbeams '' primary max_snr beam
// tab-separated with the default stream precision.
#include <arpa/inet.h>
#include <netdb.h>
#include <netinet/in.h>
#include <signal.h>
#include <sys/select.h>
#include <sys/socket.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace {

volatile sig_atomic_t g_quit = 0;
void on_signal(int) { g_quit = 1; }

// One candidate line (Candidates.C:29-50: the fields in this order, parsed
// with these types).
struct Cand {
  float snr = 0;
  int64_t sample_idx = 0;
  float sample_time = 0;
  unsigned filter = 0, dm_trial = 0;
  float dm = 0;
  unsigned members = 0;
  int64_t begin = 0, end = 0;
  unsigned beam = 0;
  unsigned nbeams = 1, primary_beam = 0;
  float max_snr = 0;

  bool parse(const std::string& line) {
    std::istringstream iss(line);
    iss >> snr >> sample_idx >> sample_time >> filter >> dm_trial >> dm >> members >> begin >>
        end >> beam;
    if (iss.fail()) return false;
    nbeams = 1;
    primary_beam = beam;
    max_snr = snr;
    iss >> std::ws;
    if (!iss.eof()) std::cout << "Candiate::Candidate too many params on input line [" << line
                              << "]" << std::endl;
    return true;
  }

  bool coincident(const Cand& c, unsigned sep_time, unsigned sep_filter) const {
    const int64_t tol = (int64_t)((float)sep_time * powf(2.f, (float)std::max(c.filter, filter)));
    const float sep_snr = 0.4f;
    return std::llabs(c.sample_idx - sample_idx) <= tol &&
           (unsigned)std::abs((int)c.dm_trial - (int)dm_trial) <= 9999u &&
           (unsigned)std::abs((int)c.filter - (int)filter) <= sep_filter &&
           (std::fabs(c.snr - snr) / (c.snr + snr)) <= sep_snr;
  }
};

std::ostream& operator<<(std::ostream& os, const Cand& c) {
  os << c.snr << "\t" << c.sample_idx << "\t" << c.sample_time << "\t" << c.filter << "\t"
     << c.dm_trial << "\t" << c.dm << "\t" << c.members << "\t" << c.begin << "\t" << c.end
     << "\t" << c.nbeams << "\t" << "\t" << c.primary_beam << "\t" << c.max_snr << "\t"
     << c.beam;
  return os;
}

// UTC "YYYY-MM-DD-hh:mm:ss" as seconds since the epoch (Candidates.C:333-355)
time_t utc_seconds(const std::string& s) {
  struct tm t;
  std::memset(&t, 0, sizeof(t));
  if (!strptime(s.c_str(), "%Y-%m-%d-%H:%M:%S", &t)) return 0;
  return timegm(&t);
}

// The candidates of one gulp of every beam (CandidateChunk).
struct Chunk {
  std::vector<std::vector<Cand>> beams;
  unsigned nbeam = 0;            // beams added so far
  bool configured = false;
  uint64_t first_sample = 0;
  std::string first_sample_utc, utc_start;

  explicit Chunk(unsigned nbeams) : beams(nbeams) {}

  void add(const Cand& c) {
    if (c.beam >= 1 && c.beam <= beams.size()) beams[c.beam - 1].push_back(c);
    else std::cerr << "coincidencer: candidate of beam " << c.beam << " outside 1.."
                   << beams.size() << " ignored" << std::endl;
  }

  // CandidateChunk::addBeam (Candidates.C:159-204)
  void add_message(const std::string& utc, const std::string& first_utc, uint64_t first,
                   unsigned nb, uint64_t n_events, std::istringstream& ss) {
    if (!configured) {
      first_sample = first;
      first_sample_utc = first_utc;
      utc_start = utc;
      configured = true;
    } else {
      if (first_sample != first)
        std::cerr << "CandidateChunk::addBeam sample mismatch [" << first_sample << " != "
                  << first << "]" << std::endl;
      if (utc_start != utc) std::cerr << "CandidateChunk::addBeam utc_start mismatch" << std::endl;
    }
    nbeam += nb;
    std::string line;
    for (uint64_t e = 0; e < n_events; ++e) {
      // a last line with no newline is dropped, as ss.getline + !ss.eof() does
      if (!std::getline(ss, line) || ss.eof()) break;
      if (line.size() > 10) {
        Cand c;
        if (c.parse(line)) add(c);
      }
    }
  }

  // CandidateChunk::compute_coincidence (Candidates.C:257-301)
  void compute(unsigned sep_time, unsigned sep_filter) {
    for (size_t i = 0; i < beams.size(); ++i)
      for (Cand& cj : beams[i]) {
        float max_snr_j = cj.snr;
        for (size_t k = 0; k < beams.size(); ++k) {
          if (k == i) continue;
          for (const Cand& cl : beams[k])
            if (cj.coincident(cl, sep_time, sep_filter)) {
              cj.nbeams++;
              if (cl.snr > max_snr_j) {
                cj.primary_beam = cl.beam;
                cj.max_snr = cl.snr;
                max_snr_j = cl.snr;
              }
              break;  // the first coincident candidate of each beam
            }
        }
      }
  }

  // CandidateChunk::write_coincident_candidates (Candidates.C:303-330)
  void write() const {
    std::string name;
    if (utc_start.empty()) {
      name = first_sample_utc + "_all.cand";
    } else {
      struct stat st;
      if (stat(utc_start.c_str(), &st) == 0 && S_ISDIR(st.st_mode)) {
        name = utc_start + "/" + first_sample_utc + "_all.cand";
      } else {
        std::cerr << "directory [" << utc_start << "] did not exist, not writing candidate file"
                  << std::endl;
        return;
      }
    }
    std::ofstream ofs(name.c_str(), std::ios::out);
    for (const auto& b : beams)
      for (const Cand& c : b) ofs << c << std::endl;
  }

  time_t relative_age(const std::string& utc) const {
    return utc_seconds(utc) - utc_seconds(first_sample_utc);
  }
};

void usage() {
  std::cout << "coincidencer [options] [beam_candidate_files]\n"
               "  ** assumes beams are counted from 1\n"
               "  -f sep_filter  separation filter [default 4]\n"
               "  -h             print this help text\n"
               "  -t sep_time    separation time [default 3]\n"
               "  -a address     interface for candidate events\n"
               "  -n nbeams      number of beams to expect data from\n"
               "  -p port        port for candidate events\n"
               "  -v             verbose output\n";
}

int listen_on(const char* address, int port) {
  const int fd = socket(AF_INET, SOCK_STREAM, 0);
  if (fd < 0) return -1;
  const int one = 1;
  setsockopt(fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one));
  sockaddr_in a{};
  a.sin_family = AF_INET;
  a.sin_port = htons((uint16_t)port);
  if (std::strcmp(address, "any") == 0) {
    a.sin_addr.s_addr = htonl(INADDR_ANY);
  } else {
    addrinfo hints{}, *res = nullptr;
    hints.ai_family = AF_INET;
    if (getaddrinfo(address, nullptr, &hints, &res) != 0 || !res) {
      close(fd);
      return -1;
    }
    a.sin_addr = reinterpret_cast<sockaddr_in*>(res->ai_addr)->sin_addr;
    freeaddrinfo(res);
  }
  if (bind(fd, reinterpret_cast<sockaddr*>(&a), sizeof(a)) != 0 || listen(fd, 64) != 0) {
    close(fd);
    return -1;
  }
  return fd;
}

// Waits up to `sec` for a connection; true when one is waiting.
bool poll_accept(int fd, double sec) {
  fd_set r;
  FD_ZERO(&r);
  FD_SET(fd, &r);
  timeval tv{(time_t)sec, (suseconds_t)((sec - (time_t)sec) * 1e6)};
  return select(fd + 1, &r, nullptr, nullptr, &tv) > 0;
}

std::string read_all(int fd) {
  std::string out;
  char buf[65536];
  for (;;) {
    const ssize_t n = recv(fd, buf, sizeof(buf), 0);
    if (n <= 0) break;
    out.append(buf, (size_t)n);
  }
  return out;
}

std::string now_str() {
  char b[64];
  const time_t t = time(nullptr);
  std::strftime(b, 64, "%Y-%m-%d-%H:%M:%S", localtime(&t));
  return b;
}

}  // namespace

int main(int argc, char* argv[]) {
  int port = 0;
  const char* address = "any";
  unsigned total_beams = 1, sep_time = 3, sep_filter = 4, verbose = 0;
  const unsigned max_chunks_to_wait = 4;
  int opt;
  while ((opt = getopt(argc, argv, "a:f:hn:p:t:v")) != -1) {
    switch (opt) {
      case 'a': address = optarg; break;
      case 't': sep_time = (unsigned)std::atoi(optarg); break;
      case 'f': sep_filter = (unsigned)std::atoi(optarg); break;
      case 'h': usage(); return 0;
      case 'n': total_beams = (unsigned)std::atoi(optarg); break;
      case 'p': port = std::atoi(optarg); break;
      case 'v': ++verbose; break;
      default: usage(); return 1;
    }
  }
  const int nfiles = argc - optind;
  if (nfiles > 0) {
    // file mode (coincidencer.C:124-142, Candidates.C:103-141)
    Chunk chunk(total_beams);
    std::string first(argv[optind]);
    chunk.first_sample_utc = first.substr(0, first.find('_'));
    chunk.nbeam = total_beams;
    chunk.configured = true;
    for (int f = 0; f < nfiles; ++f) {
      std::ifstream ifs(argv[optind + f]);
      std::string line;
      while (std::getline(ifs, line) && !ifs.eof()) {  // (Candidates.C:129-139)
        Cand c;
        if (c.parse(line)) chunk.add(c);
      }
    }
    chunk.compute(sep_time, sep_filter);
    chunk.write();
    return 0;
  }
  if (port <= 0) {
    std::cerr << "ERROR!" << std::endl;
    return 1;
  }
  // network mode (coincidencer.C:144-358)
  signal(SIGINT, on_signal);
  signal(SIGTERM, on_signal);
  const int server = listen_on(address, port);
  if (server < 0) {
    std::cerr << "coincidencer: cannot listen on " << address << ":" << port << std::endl;
    return 1;
  }
  std::cerr << "[" << now_str() << "] listening for connections on " << address << ":" << port
            << std::endl;
  std::vector<std::unique_ptr<Chunk>> chunks;
  std::string curr_first_sample_utc, prev_utc_start;
  int curr_chunk = 0;
  while (!g_quit) {
    bool waiting = true;
    while (waiting && !g_quit) {
      if (!poll_accept(server, 0.05)) continue;
      const int c = accept(server, nullptr, nullptr);
      if (c < 0) continue;
      const std::string data = read_all(c);
      close(c);
      std::istringstream iss(data);
      std::string utc_start, first_sample_utc;
      uint64_t first_sample = 0, n_events = 0;
      int first_beam = 0, nbeams = 0;
      iss >> utc_start >> first_sample_utc >> first_sample >> first_beam >> nbeams >> n_events >>
          std::ws;
      if (iss.fail()) {
        std::cerr << "caught socket exception!" << std::endl;
        continue;
      }
      if (verbose)
        std::cerr << "main: UTC_START=" << utc_start << " SAMPLE_UTC=" << first_sample_utc
                  << " SAMPLE_IDX=" << first_sample << " FIRST_BEAM=" << first_beam
                  << " NBEAMS=" << nbeams << " NUM_EVENTS=" << n_events << std::endl;
      if (prev_utc_start != utc_start) {
        std::cerr << "[" << now_str() << "] processing UTC_START=[" << utc_start << "] PREV=["
                  << prev_utc_start << "]" << std::endl;
        prev_utc_start = utc_start;
      }
      if (curr_first_sample_utc.empty() || curr_first_sample_utc != first_sample_utc) {
        curr_chunk = -1;
        time_t youngest = 0;
        for (size_t k = 0; k < chunks.size(); ++k) {
          const time_t age = chunks[k]->relative_age(first_sample_utc);
          if (age == 0) curr_chunk = (int)k;
          if (age < youngest) youngest = age;
        }
        if (curr_chunk == -1 && youngest < 0) curr_chunk = -2;  // older than every chunk
        if (curr_chunk == -1) {
          chunks.emplace_back(new Chunk(total_beams));
          curr_chunk = (int)chunks.size() - 1;
        } else if (curr_chunk == -2) {
          std::cerr << "main: new beam " << first_sample_utc << " [" << first_beam << "-"
                    << first_beam + nbeams << "] arrived too late [curr_chunk=" << curr_chunk
                    << "]" << std::endl;
        }
      }
      if (curr_chunk >= 0) {
        curr_first_sample_utc = first_sample_utc;
        Chunk& ch = *chunks[curr_chunk];
        ch.add_message(utc_start, first_sample_utc, first_sample, (unsigned)nbeams, n_events, iss);
        if (ch.nbeam == total_beams || chunks.size() > max_chunks_to_wait) waiting = false;
      }
    }
    if (g_quit) break;
    if (chunks.size() > max_chunks_to_wait) curr_chunk = 0;
    Chunk& ch = *chunks[curr_chunk];
    ch.compute(sep_time, sep_filter);
    ch.write();
    std::cerr << "[" << now_str() << "] wrote cands " << ch.first_sample_utc << std::endl;
    chunks.erase(chunks.begin() + curr_chunk);
    curr_chunk = 0;
    curr_first_sample_utc.clear();
  }
  close(server);
  return 0;
}
