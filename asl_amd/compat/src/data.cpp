// Data containers, acl globals and the parameters manager of the compat layer.
#include <fstream>
#include <iostream>

#include <acl/aclTypes.h>
#include <aslDataInc.h>
#include <data/aslDataWithGhostNodes.h>
#include <utilities/aslParametersManager.h>
#include <writers/aslWriter.h>

namespace acl {
Hardware hardware;
}

namespace asl {

Writer *Writer::current = nullptr;
ParametersManager *ParametersManager::current = nullptr;

std::vector<float> AbstractData::componentAsFloat(unsigned int c) {
  const std::vector<double> &v = component(c);
  return std::vector<float>(v.begin(), v.end());
}

void initData(SPAbstractData d, double a) {
  if (!d) errorMessage("initData: the container is not initialized");
  for (auto &c : d->values_) std::fill(c.begin(), c.end(), a);
}

void initData(SPAbstractData d, const AVec<double> &a) {
  if (!d) errorMessage("initData: the container is not initialized");
  if (a.getSize() != d->getNComponents())
    errorMessage("initData: the vector size differs from the number of components");
  for (unsigned int c = 0; c < a.getSize(); ++c) std::fill(d->values_[c].begin(), d->values_[c].end(), a[c]);
}

SPDataWithGhostNodesACLData generateDataContainerACL_SP(const Block &b, acl::TypeID t, unsigned int n,
                                                        unsigned int gN) {
  return std::make_shared<DataWithGhostNodesACLData>(b, n, t, int(gN));
}

void ParametersManager::set(const std::string &key, const std::string &value) {
  auto it = entries_.find(key);
  if (it == entries_.end()) errorMessage("ParametersManager: unknown parameter \"" + key + "\"");
  it->second.assign(value);
  it->second.set = true;
}

void ParametersManager::load(std::string paramFile) {
  std::ifstream in(paramFile);
  if (!in) errorMessage("ParametersManager: cannot open parameters file " + paramFile);
  const size_t slash = paramFile.find_last_of('/');
  dir_ = slash == std::string::npos ? "./" : paramFile.substr(0, slash + 1);
  std::string line;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    auto trim = [](std::string s) {
      const size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t\r");
      return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    };
    set(trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
  }
  check_required();
}

void ParametersManager::check_required() {
  for (auto &e : entries_)
    if (e.second.required && !e.second.set) errorMessage("ParametersManager: parameter \"" + e.first + "\" is required");
}

void ParametersManager::print_help(const std::string &name) const {
  std::cout << name << " options:\n  -h [ --help ]  this message\n  --platform, --device  accepted (one GPU)\n";
  for (auto &e : entries_)
    std::cout << "  --" << e.first << "  " << e.second.description
              << (e.second.units.empty() ? "" : " [" + e.second.units + "]") << "\n";
}

ApplicationParametersManager::ApplicationParametersManager(const char *applicationName_,
                                                           const char *applicationVersion_)
    : applicationName(applicationName_), applicationVersion(applicationVersion_) {
  add(platform, std::string("default"), "platform", "platform (ignored: one GPU)", "");
  add(device, std::string("default"), "device", "device (ignored: one GPU)", "");
}

void ApplicationParametersManager::load(int argc, char *argv[]) {
  std::string file;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      print_help(applicationName + " " + applicationVersion);
      std::exit(0);
    }
    if (a.rfind("--", 0) == 0) {
      a = a.substr(2);
      const size_t eq = a.find('=');
      if (eq != std::string::npos) {
        set(a.substr(0, eq), a.substr(eq + 1));
      } else {
        if (i + 1 >= argc) errorMessage("ParametersManager: option --" + a + " needs a value");
        set(a, argv[++i]);
      }
    } else {
      file = a;
    }
  }
  if (!file.empty())
    ParametersManager::load(file);
  else
    check_required();
}

}  // namespace asl
