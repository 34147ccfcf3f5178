// Test support: generate and hipRTC-compile the engine's kernel sources for
// gfx950 without a GPU, from form descriptions written by
// tests/test_codegen_cpu.py (rank, shapes, coefficient layout and each
// integral's UFCx body + staged description).  Built by that test against
// the library's object files.
//
//   codegen_check form  <description> <out.co>        one form's translation unit
//   codegen_check batch <out.co> <description>...      the one-launch batch of several forms
//   codegen_check system <out.co> <J> <F> [<E>]        the fused Jacobian/residual/functional launch
//
// Prints "OK <bytes>" or "FAIL <message>"; the test reads the kernels'
// resource use from the code object's metadata.
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "mf_internal.h"

namespace {

struct Loaded {
  mf_mesh mesh;
  mf_dofmap dm0, dm1;
  mf_form form;
};

std::unique_ptr<Loaded> load(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  auto L = std::make_unique<Loaded>();
  int rank, nv, nf, W, nd0, bs0, nd1, bs1, nint, ncoef;
  in >> rank >> nv >> nf >> W >> nd0 >> bs0 >> nd1 >> bs1 >> nint >> ncoef;
  L->mesh.nv = nv;
  L->mesh.nf = nf;
  L->dm0.nd = nd0;
  L->dm0.bs = bs0;
  L->dm1.nd = nd1;
  L->dm1.bs = bs1;
  mf_form& f = L->form;
  f.rank = rank;
  f.mesh = &L->mesh;
  f.dm[0] = rank >= 1 ? &L->dm0 : nullptr;
  f.dm[1] = rank == 2 ? (nd0 == nd1 && bs0 == bs1 ? &L->dm0 : &L->dm1) : nullptr;
  f.W = W;
  for (int c = 0; c < ncoef; ++c) {
    mf::Coef co;
    in >> co.offset >> co.width;
    f.coefs.push_back(std::move(co));
  }
  for (int i = 0; i < nint; ++i) {
    auto I = std::make_unique<mf::Integral>();
    size_t ls = 0, ss = 0;
    in >> I->type >> I->index >> I->tab_name >> ls >> ss;
    in.get();
    I->tab_src.resize(ls);
    in.read(&I->tab_src[0], static_cast<std::streamsize>(ls));
    I->staged_src.resize(ss);
    if (ss) in.read(&I->staged_src[0], static_cast<std::streamsize>(ss));
    f.ints.push_back(std::move(I));
  }
  if (!in) throw std::runtime_error("truncated description " + path);
  return L;
}

void write(const std::string& path, const std::vector<char>& code) {
  std::ofstream(path, std::ios::binary).write(code.data(), static_cast<std::streamsize>(code.size()));
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    const auto opts = mf::rtc_compile_options("gfx950");
    if (mode == "form" && argc == 4) {
      auto L = load(argv[2]);
      const std::string src = mf::compile_source(L->form);
      auto code = mf::rtc_compile(src, "codegen_check", opts, "", nullptr);
      write(argv[3], code);
      std::cout << "OK " << code.size() << "\n";
      return 0;
    }
    if (mode == "batch" && argc >= 4) {
      std::vector<std::unique_ptr<Loaded>> Ls;
      std::vector<const mf_form*> fs;
      std::vector<const mf::Integral*> is;
      for (int k = 3; k < argc; ++k) {
        Ls.push_back(load(argv[k]));
        mf::compile_source(Ls.back()->form);      // defaults the owner-kernel shape options
        fs.push_back(&Ls.back()->form);
        is.push_back(Ls.back()->form.ints[0].get());
      }
      int64_t lds = 0;
      const std::string src = mf::generate_batch_source(fs, is, 256, &lds);
      auto code = mf::rtc_compile(src, "codegen_check", opts, "", nullptr);
      write(argv[2], code);
      std::cout << "OK " << code.size() << " lds " << lds << "\n";
      return 0;
    }
    if (mode == "system" && (argc == 5 || argc == 6)) {
      auto J = load(argv[3]);
      auto F = load(argv[4]);
      std::unique_ptr<Loaded> E = argc == 6 ? load(argv[5]) : nullptr;
      const std::string src =
          mf::generate_system_source(J->form, F->form, E ? &E->form : nullptr, 256, 0);
      auto code = mf::rtc_compile(src, "codegen_check", opts, "", nullptr);
      write(argv[2], code);
      // constant tables: declared inside the bodies, shared file-scope arrays
      // in the kernel source, function-local ones left in it
      auto count = [](const std::string& t, const std::string& key) {
        size_t n = 0;
        for (size_t p = t.find(key); p != std::string::npos; p = t.find(key, p + key.size())) ++n;
        return n;
      };
      size_t declared = 0;
      for (const mf_form* f : {&J->form, &F->form, E ? &E->form : nullptr})
        if (f)
          for (auto& I : f->ints) declared += count(I->tab_src, "static const double ");
      std::cout << "OK " << code.size() << " tables " << declared << " "
                << count(src, "static __constant__ const double mf_tab_") << " "
                << count(src, "static const double ") << "\n";
      return 0;
    }
    std::cout << "FAIL usage: codegen_check form <description> <out.co> | batch <out.co> "
                 "<description>... | system <out.co> <J> <F> [<E>]\n";
    return 2;
  } catch (const std::exception& e) {
    std::cout << "FAIL " << e.what() << "\n";
    return 1;
  }
}
