// loquat_io.hpp -- case/IO surface of the reference (loquat_io.h:39-97).
//
// Reads <project>.xml (typed <parameters> children), <project>.dat (boundary
// rows with 31 columns, then soil rows with 27) and <project>.motion, with the
// same key names, column order and fatal-error behaviour as the reference
// (loquat_io.cu:62-312).  Writes output_<N>.xml, <N>.dat (restartable, same
// format as the input) or <id>_<N>.vtu files (output_format 2), and
// time_history.txt.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "retainwall.h"

class LoquatIo {
 public:
  explicit LoquatIo(const std::string& project_name);

  void LoadParameters();  // loquat_io.cu:75-202
  void LoadParticles();   // loquat_io.cu:205-312
  void SaveXml();         // loquat_io.cu:315-399
  void SaveParticles();   // loquat_io.cu:430-680 (.dat for format 1, .vtu for format 2)

  lq_params* GetParameters() { return &par_; }
  std::vector<int32_t>& GetCodes() { return codes_; }
  std::vector<float>& GetPosPres() { return pospres_; }
  std::vector<float>& GetVelRhop() { return velrhop_; }
  std::vector<float>& GetIdMassIdx() { return idmassidx_; }
  std::vector<float>& GetStr() { return str_; }
  std::vector<float>& GetStn() { return stn_; }
  std::vector<float>& GetNormalMaterial() { return normal_; }
  std::vector<float>& GetAccad() { return accad_; }
  std::vector<lq_motion_table>& GetMotionTables() { return motion_; }
  const std::string& ProjectName() const { return project_name_; }

 private:
  struct Value {
    std::string type;
    std::string text;
  };
  const Value& CheckParameter(const std::string& name) const;  // loquat_io.cu:62-72
  int GetInt(const std::string& name) const;
  float GetScalar(const std::string& name) const;
  bool GetBool(const std::string& name) const;
  void SaveVtu();  // loquat_io.cu:470-680

  std::string project_name_;
  std::map<std::string, Value> xml_;
  lq_params par_{};
  std::vector<int32_t> codes_;
  std::vector<float> pospres_, velrhop_, idmassidx_, str_, stn_, normal_, accad_;
  std::vector<lq_motion_table> motion_;
};
