// esf_cfg.h -- host-side cfg reader and scene helpers shared by libesf.so's
// scene setup and the headless driver.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "../../include/esf.h"

namespace esf {

// key/value store with the reference's SIMPLE_PARSER semantics
struct Cfg {
  std::map<std::string, std::string> vals;
  bool parse(const std::string& path, std::string* err);
  bool has(const std::string& k) const;
  double getFloat(const std::string& k, double def) const;
  int getInt(const std::string& k, int def) const;
  std::string getString(const std::string& k, const std::string& def) const;
  void getVec3(const std::string& k, const double def[3], double out[3]) const;
};

// a FIELD_3Df in rest space (the solid's rest mass density)
struct RestField {
  int X = 0, Y = 0, Z = 0;
  double dh = 0, origin[3] = {0, 0, 0};
  std::vector<double> data;
};

void paramsFromCfg(const Cfg& c, double dh, esf_params* p);
bool readMassGz(const std::string& path, RestField* f);
void rasterizeSphere(double radius, double massDh, double density, RestField* f);
void rotationFromCfg(const double r[3], double R[9]);

}  // namespace esf
