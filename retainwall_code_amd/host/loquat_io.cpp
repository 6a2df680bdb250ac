// loquat_io.cpp -- XML / .dat / .motion input and plain-text output.
#include "loquat_io.hpp"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace {

[[noreturn]] void die() { std::exit(EXIT_FAILURE); }

// Minimal reader for the reference's parameter files: an optional prolog,
// one root element, and children <name type="...">text</name>.
bool parse_parameter_xml(const std::string& s, std::map<std::string, std::pair<std::string, std::string>>& out) {
  size_t p = 0;
  auto skip_ws = [&]() { while (p < s.size() && std::isspace((unsigned char)s[p])) p++; };
  auto skip_misc = [&]() {
    for (;;) {
      skip_ws();
      if (s.compare(p, 4, "<!--") == 0) {
        size_t e = s.find("-->", p + 4);
        if (e == std::string::npos) return false;
        p = e + 3;
      } else if (s.compare(p, 2, "<?") == 0) {
        size_t e = s.find("?>", p + 2);
        if (e == std::string::npos) return false;
        p = e + 2;
      } else {
        return true;
      }
    }
  };
  auto read_name = [&]() {
    size_t b = p;
    while (p < s.size() && !std::isspace((unsigned char)s[p]) && s[p] != '>' && s[p] != '/') p++;
    return s.substr(b, p - b);
  };
  if (!skip_misc() || p >= s.size() || s[p] != '<') return false;
  p++;
  std::string root = read_name();
  if (root.empty()) return false;
  size_t gt = s.find('>', p);
  if (gt == std::string::npos) return false;
  p = gt + 1;
  for (;;) {
    if (!skip_misc()) return false;
    if (p >= s.size()) return false;
    if (s.compare(p, 2, "</") == 0) return true;  // end of root
    if (s[p] != '<') return false;
    p++;
    std::string name = read_name();
    if (name.empty()) return false;
    std::string type;
    bool empty = false;
    for (;;) {  // attributes
      skip_ws();
      if (p >= s.size()) return false;
      if (s[p] == '/') { empty = true; p = s.find('>', p); if (p == std::string::npos) return false; p++; break; }
      if (s[p] == '>') { p++; break; }
      size_t eq = s.find('=', p);
      if (eq == std::string::npos) return false;
      std::string an = s.substr(p, eq - p);
      while (!an.empty() && std::isspace((unsigned char)an.back())) an.pop_back();
      p = eq + 1;
      skip_ws();
      if (p >= s.size() || (s[p] != '"' && s[p] != '\'')) return false;
      char q = s[p++];
      size_t e = s.find(q, p);
      if (e == std::string::npos) return false;
      if (an == "type") type = s.substr(p, e - p);
      p = e + 1;
    }
    std::string text;
    if (!empty) {
      size_t e = s.find("</", p);
      if (e == std::string::npos) return false;
      text = s.substr(p, e - p);
      p = e + 2;
      std::string close = read_name();
      if (close != name) return false;
      p = s.find('>', p);
      if (p == std::string::npos) return false;
      p++;
    }
    out[name] = {type, text};
  }
}

}  // namespace

LoquatIo::LoquatIo(const std::string& project_name) : project_name_(project_name) {
  std::memset(&par_, 0, sizeof(par_));
}

const LoquatIo::Value& LoquatIo::CheckParameter(const std::string& name) const {
  auto it = xml_.find(name);
  if (it == xml_.end()) {
    std::cout << "Error! " << name << " does not exist in the xml file!" << std::endl;
    die();
  }
  return it->second;
}

int LoquatIo::GetInt(const std::string& name) const {
  const Value& v = CheckParameter(name);
  return std::atoi(v.text.c_str());
}
float LoquatIo::GetScalar(const std::string& name) const {
  const Value& v = CheckParameter(name);
  return float(std::atof(v.text.c_str()));
}
bool LoquatIo::GetBool(const std::string& name) const {
  const Value& v = CheckParameter(name);
  return bool(std::atof(v.text.c_str()));
}

void LoquatIo::LoadParameters() {
  std::cout << std::endl << "Loading parameters ..." << std::endl;
  std::string xml_name = project_name_ + ".xml";
  std::ifstream f(xml_name);
  std::stringstream buf;
  if (f) buf << f.rdbuf();
  std::map<std::string, std::pair<std::string, std::string>> raw;
  if (!f || !parse_parameter_xml(buf.str(), raw)) {
    std::cout << "Error! Cannot load the parameter file: " << xml_name << std::endl;
    std::cout << "Check the XML file, the names much match!" << std::endl;
    die();
  }
  std::cout << "The parameter file \"" << xml_name << "\" is loaded sucessfully" << std::endl;
  for (auto& kv : raw) {
    const std::string& t = kv.second.first;
    if (t != "int" && t != "scalar" && t != "model_mat" && t != "dimension" && t != "bool") {
      std::cout << "Wrong data type for " << kv.first << "!!!" << std::endl;  // loquat_io.cu:120-123
      die();
    }
    xml_[kv.first] = {t, kv.second.second};
  }
  lq_params& p = par_;
  std::memset(&p, 0, sizeof(p));
  p.dim = GetInt("dimension-of-problem");
  p.n_motion = GetInt("number-of-motion");
  p.periodbound = GetBool("period-bound");
  if (p.periodbound) {
    const char* k[4] = {"A", "B", "C", "D"};
    for (int i = 0; i < 4; i++) {
      p.planeleft[i] = GetScalar(std::string("planeleft-") + k[i]);
      p.planeright[i] = GetScalar(std::string("planeright-") + k[i]);
    }
  }
  p.np = GetInt("number-of-particles");
  p.npb = GetInt("number-of-boundary-particles");
  p.npm = GetInt("number-of-material-particles");
  if (p.np != p.npb + p.npm) {
    std::cout << "Error! Wrong number of particles ..." << std::endl;
    die();
  }
  const char* ax[3] = {"x", "y", "z"};
  for (int i = 0; i < 3; i++) {
    p.domain_size[i] = GetScalar(std::string("domain-size-") + ax[i]);
    p.domain_min[i] = GetScalar(std::string("domain-min-") + ax[i]);
  }
  p.dr = GetScalar("particle-size");
  p.hdr = GetScalar("h-over-dr");
  for (int i = 0; i < 3; i++) p.gravity[i] = GetScalar(std::string("acceleration-") + ax[i]);
  p.rho0_m = GetScalar("reference-density-material");
  p.damping = GetScalar("damping");
  p.n_model = GetInt("number-of-model");
  if (p.n_model < 0 || p.n_model > LQ_MODEL_NUMBER_MAX) {
    std::cout << "Error! number-of-model out of range" << std::endl;
    die();
  }
  for (int i = 0; i < p.n_model; i++)
    for (int j = 0; j < LQ_MATERIAL_SIZE_MAX; j++)
      p.model_par[i][j] = GetScalar("model-" + std::to_string(i) + "-par-" + std::to_string(j));
  p.stress_shepard_filter = GetInt("stress-shepard-filter");
  p.dynamic_time_step = GetInt("dynamic-time-step");
  p.integration_method = GetInt("integration-method");
  p.viscosity_type = GetInt("viscosity-type");
  if (p.viscosity_type == 1) p.artificial_viscosity_alpha_m = GetScalar("artificial-viscosity-alpha-material");
  p.cs_m = GetScalar("speed-of-sound-material");
  p.cfl = GetScalar("cfl-number");
  p.artificial_pressure_coefficient_material = GetScalar("artificial-pressure-coefficient-material");
  p.time_max = GetScalar("time-max");
  p.time = GetScalar("time");
  p.dt = GetScalar("time-step");
  p.output_frequency = GetScalar("output-frequency");
  p.output_number = GetInt("output-number");
  p.output_format = GetInt("output-format");
  std::cout << "Finish reading parameters" << std::endl << std::endl;
}

void LoquatIo::LoadParticles() {
  std::cout << "Loading particle data ..." << std::endl;
  const int np = par_.np, npb = par_.npb, npm = par_.npm;
  codes_.assign(size_t(np) * 4, 0);
  pospres_.assign(size_t(np) * 4, 0.f);
  velrhop_.assign(size_t(np) * 4, 0.f);
  idmassidx_.assign(size_t(np) * 4, 0.f);
  str_.assign(size_t(np) * 6, 0.f);
  stn_.assign(size_t(np) * 6, 0.f);
  normal_.assign(size_t(std::max(npb, 0)) * 4, 0.f);
  accad_.assign(size_t(np) * 4, 0.f);
  std::string dat = project_name_ + ".dat";
  std::ifstream in(dat);
  if (!in) {
    std::cout << "Error! Cannot open " << dat << " !" << std::endl;
    die();
  }
  for (int i = 0; i < npb + npm; i++) {
    int32_t* c = &codes_[4 * size_t(i)];
    float* p = &pospres_[4 * size_t(i)];
    float* v = &velrhop_[4 * size_t(i)];
    float* m = &idmassidx_[4 * size_t(i)];
    float* s = &str_[6 * size_t(i)];
    float* e = &stn_[6 * size_t(i)];
    in >> c[0] >> c[1] >> c[2] >> c[3] >> p[0] >> p[1] >> p[2] >> p[3] >> v[0] >> v[1] >> v[2] >> v[3] >> m[0] >>
        m[1] >> m[3] >> s[0] >> s[1] >> s[2] >> s[3] >> s[4] >> s[5] >> e[0] >> e[1] >> e[2] >> e[3] >> e[4] >> e[5];
    if (i < npb) {
      float* n = &normal_[4 * size_t(i)];
      in >> n[0] >> n[1] >> n[2] >> n[3];
    }
    m[2] = float(i);  // loquat_io.cu:240,253
    if (!in) {
      std::cout << "Error! Cannot read particle " << i << " of " << dat << " !" << std::endl;
      die();
    }
  }
  std::cout << "Finish reading boundary and soil particles ..." << std::endl;
  if (par_.n_motion > 0) {
    motion_.assign(size_t(par_.n_motion), lq_motion_table{});
    std::string mf = project_name_ + ".motion";
    std::ifstream ms(mf);
    if (!ms) {
      std::cout << "Error! Cannot open " << mf << " !" << std::endl;
      die();
    }
    for (int i = 0; i < par_.n_motion; i++) {
      lq_motion_table& t = motion_[size_t(i)];
      std::memset(&t, 0, sizeof(t));
      ms >> t.is_translation >> t.is_rotation;
      if (t.is_translation) {
        ms >> t.table_size_trans;
        for (int j = 0; j < t.table_size_trans && j < LQ_TABLE_MAX; j++)
          ms >> t.trans_time[j] >> t.trans_vel[j][0] >> t.trans_vel[j][1] >> t.trans_vel[j][2] >> t.trans_acc[j][0] >>
              t.trans_acc[j][1] >> t.trans_acc[j][2];
      }
      if (t.is_rotation) {
        ms >> t.pattern >> t.pos_rp[0] >> t.pos_rp[1] >> t.pos_rp[2] >> t.rotat_vel[0] >> t.rotat_vel[1] >>
            t.rotat_vel[2] >> t.table_size_rotat;
        for (int j = 0; j < t.table_size_rotat && j < LQ_TABLE_MAX; j++)
          ms >> t.rotat_time[j] >> t.rotat_acc[j][0] >> t.rotat_acc[j][1] >> t.rotat_acc[j][2];
      }
      if (!ms) {
        std::cout << "Error! Cannot read motion " << i << " of " << mf << " !" << std::endl;
        die();
      }
      // loquat_io.cu:289-302: the initial imposed motion
      lq_rigid_motion& mv = par_.mv[i];
      mv.is_trans = t.is_translation;
      mv.is_rotat = t.is_rotation;
      for (int k = 0; k < 3; k++) mv.trans_vel[k] = t.is_translation ? t.trans_vel[0][k] : 0.f;
      if (t.is_rotation) {
        mv.pattern = t.pattern;
        for (int k = 0; k < 3; k++) { mv.pos_rp[k] = t.pos_rp[k]; mv.rotat_vel[k] = t.rotat_vel[k]; }
      }
    }
    std::cout << "Finish loading motion infomation ..." << std::endl;
  }
}

void LoquatIo::SaveXml() {
  std::ostringstream name;
  name << project_name_ << "/output_" << par_.output_number << ".xml";
  std::ofstream o(name.str());
  if (!o) {
    std::cout << "Error! Cannot write file " << name.str() << " !" << std::endl;
    return;
  }
  const lq_params& p = par_;
  auto S = [&](const char* k, double v) { o << "  <" << k << " type=\"scalar\">" << float(v) << "</" << k << ">\n"; };
  auto I = [&](const char* k, int v, const char* t = "int") { o << "  <" << k << " type=\"" << t << "\">" << v << "</" << k << ">\n"; };
  o << "<?xml version=\"1.0\" encoding=\"utf-8\" ?>\n<parameters>\n";
  I("dimension-of-problem", p.dim, "dimension");
  I("number-of-motion", p.n_motion);
  I("period-bound", p.periodbound, "bool");
  if (p.periodbound) {
    const char* kl[4] = {"planeleft-A", "planeleft-B", "planeleft-C", "planeleft-D"};
    const char* kr[4] = {"planeright-A", "planeright-B", "planeright-C", "planeright-D"};
    for (int i = 0; i < 4; i++) S(kl[i], p.planeleft[i]);
    for (int i = 0; i < 4; i++) S(kr[i], p.planeright[i]);
  }
  I("number-of-particles", p.npb + p.npm);
  I("number-of-boundary-particles", p.npb);
  I("number-of-material-particles", p.npm);
  S("domain-size-x", p.domain_size[0]); S("domain-size-y", p.domain_size[1]); S("domain-size-z", p.domain_size[2]);
  S("domain-min-x", p.domain_min[0]); S("domain-min-y", p.domain_min[1]); S("domain-min-z", p.domain_min[2]);
  S("particle-size", p.dr);
  S("h-over-dr", p.hdr);
  S("acceleration-x", p.gravity[0]); S("acceleration-y", p.gravity[1]); S("acceleration-z", p.gravity[2]);
  S("reference-density-material", p.rho0_m);
  S("damping", p.damping);
  I("number-of-model", p.n_model);
  for (int i = 0; i < p.n_model; i++)
    for (int j = 0; j < LQ_MATERIAL_SIZE_MAX; j++) {
      std::string k = "model-" + std::to_string(i) + "-par-" + std::to_string(j);
      o << "  <" << k << " type=\"scalar\">" << p.model_par[i][j] << "</" << k << ">\n";
    }
  I("stress-shepard-filter", p.stress_shepard_filter);
  I("dynamic-time-step", p.dynamic_time_step);
  I("integration-method", p.integration_method);
  I("viscosity-type", p.viscosity_type);
  if (p.viscosity_type == 1) S("artificial-viscosity-alpha-material", p.artificial_viscosity_alpha_m);
  S("speed-of-sound-material", p.cs_m);
  S("cfl-number", p.cfl);
  S("artificial-pressure-coefficient-material", p.artificial_pressure_coefficient_material);
  S("time-max", p.time_max);
  S("time", p.time);
  S("time-step", p.dt);
  S("output-frequency", p.output_frequency);
  I("output-number", p.output_number);
  I("output-format", p.output_format);
  o << "</parameters>\n";
}

// output_format 1: plain .dat in the input format (boundary rows with
// normals, then soil rows), so that N.dat + output_N.xml restart a run.
// Deleted particles (slots past npb+npm) are not written.  output_format 2:
// ParaView .vtu files (SaveVtu).
void LoquatIo::SaveParticles() {
  if (par_.output_format == 2) {
    SaveVtu();
    return;
  }
  std::ostringstream name;
  name << project_name_ << "/" << par_.output_number << ".dat";
  std::ofstream o(name.str());
  if (!o) {
    std::cout << "Error! Cannot write in " << name.str() << " !" << std::endl;
    return;
  }
  o.precision(9);
  for (int i = 0; i < par_.npb + par_.npm; i++) {
    const int32_t* c = &codes_[4 * size_t(i)];
    const float* p = &pospres_[4 * size_t(i)];
    const float* v = &velrhop_[4 * size_t(i)];
    const float* m = &idmassidx_[4 * size_t(i)];
    const float* s = &str_[6 * size_t(i)];
    const float* e = &stn_[6 * size_t(i)];
    o << c[0] << " " << c[1] << " " << c[2] << " " << c[3] << " " << p[0] << " " << p[1] << " " << p[2] << " " << p[3]
      << " " << v[0] << " " << v[1] << " " << v[2] << " " << v[3] << " " << m[0] << " " << m[1] << " " << m[3];
    for (int k = 0; k < 6; k++) o << " " << s[k];
    for (int k = 0; k < 6; k++) o << " " << e[k];
    if (i < par_.npb) {
      const float* n = &normal_[4 * size_t(i)];
      o << " " << n[0] << " " << n[1] << " " << n[2] << " " << n[3];
    }
    o << "\n";
  }
}

// loquat_io.cu:470-680: one ASCII UnstructuredGrid file <id>_<N>.vtu per
// value of idmassidx.x (boundary block ids, 100 for soil, 999 for deleted),
// the slots grouped by a stable sort on that id, one vertex cell per
// particle.  Point data as the reference writes it: Material (codes.w),
// Ori_ID (idmassidx.w), Velocity, Acceleration (accad, which holds the
// accumulated displacement after a corrector), Density, Mass, Normal for
// boundary groups (id < 100), Stress and Strain in xx yy zz xy yz xz order;
// Points are the absolute positions.
void LoquatIo::SaveVtu() {
  const int np = par_.np, npb = par_.npb;
  std::vector<int> key(static_cast<size_t>(np)), order(static_cast<size_t>(np));
  for (int i = 0; i < np; i++) {
    key[size_t(i)] = static_cast<int>(idmassidx_[4 * size_t(i)]);
    order[size_t(i)] = i;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[size_t(a)] < key[size_t(b)]; });
  std::cout << "|";
  for (int g0 = 0; g0 < np;) {
    const int id = key[size_t(order[size_t(g0)])];
    int g1 = g0;
    while (g1 < np && key[size_t(order[size_t(g1)])] == id) g1++;
    const int n = g1 - g0;
    std::cout << id << ": " << n << "|";
    std::ostringstream name;
    name << project_name_ << "/" << id << "_" << par_.output_number << ".vtu";
    std::ofstream o(name.str());
    if (!o) {
      std::cout << "Error! Cannot write file " << name.str() << " !" << std::endl;
      g0 = g1;
      continue;
    }
    auto each = [&](auto&& emit) {
      for (int k = g0; k < g1; k++) emit(size_t(order[size_t(k)]));
    };
    auto vec3 = [&](const char* nm, const std::vector<float>& a, int width) {
      o << "        <DataArray type=\"Float64\" Name=\"" << nm << "\" NumberOfComponents=\"3\">\n";
      each([&](size_t i) { o << a[width * i] << " " << a[width * i + 1] << " " << a[width * i + 2] << " "; });
      o << "</DataArray>\n";
    };
    auto scalar = [&](const char* nm, auto&& value) {
      o << "        <DataArray type=\"Float64\" Name=\"" << nm << "\">\n";
      each([&](size_t i) { o << value(i) << " "; });
      o << "</DataArray>\n";
    };
    auto tensor = [&](const char* nm, const std::vector<float>& t) {
      o << "        <DataArray type=\"Float64\" Name=\"" << nm << "\" NumberOfComponents=\"6\">\n";
      each([&](size_t i) {
        const float* v = &t[6 * i];
        o << v[0] << " " << v[1] << " " << v[2] << " " << v[3] << " " << v[5] << " " << v[4] << " ";
      });
      o << "</DataArray>\n";
    };
    o << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
         "compressor=\"vtkZLibDataCompressor\">\n"
      << "  <UnstructuredGrid GhostLevel=\"0\">\n"
      << "    <Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << n << "\" Name=\"" << id << "\">\n"
      << "      <PointData>\n";
    scalar("Material", [&](size_t i) { return codes_[4 * i + 3]; });
    scalar("Ori_ID", [&](size_t i) { return idmassidx_[4 * i + 3]; });
    vec3("Velocity", velrhop_, 4);
    vec3("Acceleration", accad_, 4);
    scalar("Density", [&](size_t i) { return velrhop_[4 * i + 3]; });
    scalar("Mass", [&](size_t i) { return idmassidx_[4 * i + 1]; });
    if (id < 100) {  // MATERIAL_NUMBER, types.h:37
      o << "        <DataArray type=\"Float64\" Name=\"Normal\" NumberOfComponents=\"3\">\n";
      each([&](size_t i) {
        const bool b = i < size_t(npb);
        o << (b ? normal_[4 * i] : 0.f) << " " << (b ? normal_[4 * i + 1] : 0.f) << " " << (b ? normal_[4 * i + 2] : 0.f)
          << " ";
      });
      o << "</DataArray>\n";
    }
    tensor("Stress", str_);
    tensor("Strain", stn_);
    o << "      </PointData>\n      <CellData>\n      </CellData>\n      <Points>\n";
    vec3("Points", pospres_, 4);
    o << "      </Points>\n      <Cells>\n"
      << "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
    for (int j = 0; j < n; j++) o << j << " ";
    o << "</DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\"> \n";
    for (int j = 0; j < n; j++) o << j + 1 << " ";
    o << "</DataArray>\n        <DataArray type=\"Int64\" Name=\"types\" format=\"ascii\">\n";
    for (int j = 0; j < n; j++) o << 1 << " ";
    o << "</DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    g0 = g1;
  }
  std::cout << std::endl;
}
