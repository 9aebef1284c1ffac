// jitc.cpp — the runtime compiler of libdfrocm's shape-specialised kernels,
// run as a child process of the library (dfr_jitc beside libdfrocm.so).
//
// hiprtc compiles in the calling process; a compiler crash there would end
// the caller's process. The library therefore hands each generated source
// to this program and loads the code object it writes, so that a failed or
// crashed compile only costs the specialised kernel (the precompiled one
// runs instead).
//
//   dfr_jitc <source file> <code object file> <option>...
//
// Exit status 0: the code object was written; 1: usage or I/O error; 2: the
// source did not compile (the compiler's log goes to stderr).
#include <hip/hiprtc.h>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: dfr_jitc <source> <output> <option>...\n");
    return 1;
  }
  std::ifstream in(argv[1], std::ios::binary);
  if (!in) {
    fprintf(stderr, "dfr_jitc: cannot read %s\n", argv[1]);
    return 1;
  }
  std::stringstream buf;
  buf << in.rdbuf();
  const std::string src = buf.str();
  std::vector<const char *> opts(argv + 3, argv + argc);
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "dfr_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) return 1;
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      std::string log(n, '\0');
      if (hiprtcGetProgramLog(prog, &log[0]) == HIPRTC_SUCCESS) fprintf(stderr, "%s\n", log.c_str());
    }
    hiprtcDestroyProgram(&prog);
    return 2;
  }
  size_t size = 0;
  if (hiprtcGetCodeSize(prog, &size) != HIPRTC_SUCCESS || size == 0) return 2;
  std::vector<char> code(size);
  if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS) return 2;
  hiprtcDestroyProgram(&prog);
  std::ofstream out(argv[2], std::ios::binary);
  out.write(code.data(), (std::streamsize)code.size());
  return out ? 0 : 1;
}
