// Library-level entry points: version and named physical constants.
#include <string.h>
#include "common.h"
#include "consts.h"

extern "C" int pywrf_version(void) { return 1; }

extern "C" int pywrf_constant(const char* name, double* value) {
  struct Entry {
    const char* n;
    double v;
  };
  static const Entry table[] = {
      {"r_d", pw::r_d}, {"p1000mb", pw::p1000mb}, {"t0", pw::t0}, {"g", pw::g}, {"rcp", pw::rcp},
      {"SVP1", pw::SVP1}, {"SVP2", pw::SVP2}, {"SVP3", pw::SVP3}, {"SVPT0", pw::SVPT0}, {"EP_2", pw::EP_2},
      {"XLV", pw::XLV}, {"prandtl", pw::prandtl}, {"c_s", pw::c_s}, {"Cp", pw::Cp}, {"r_v", pw::r_v},
      {"cpovcv", pw::cpovcv}, {"cvpm", pw::cvpm}, {"reradius", pw::reradius}, {"rvovrd", pw::rvovrd}, {"Cpv", pw::Cpv},
      {"ep_2", pw::ep_2}, {"epsilon", pw::epsilon}, {"XLS", pw::XLS}, {"XLF", pw::XLF},
      {"rhoair0", pw::rhoair0}, {"rhowater", pw::rhowater}, {"cliq", pw::cliq}, {"cice", pw::cice},
      {"psat", pw::psat}, {"dtcldcr", pw::dtcldcr}, {"n0r", pw::n0r}, {"avtr", pw::avtr}, {"bvtr", pw::bvtr},
      {"bvts", pw::bvts}, {"bvtg", pw::bvtg}, {"n0s", pw::n0s}, {"n0smax", pw::n0smax}, {"n0g", pw::n0g},
      {"alpha", pw::alpha}, {"pfrz1", pw::pfrz1}, {"pfrz2", pw::pfrz2}, {"qcrmin", pw::qcrmin},
      {"dens", pw::dens}, {"qs0", pw::qs0}, {"qc0", pw::qc0}, {"qck1", pw::qck1}, {"xncr", pw::xncr},
      {"dicon", pw::dicon}, {"dimax", pw::dimax}, {"pvtr", pw::pvtr}, {"pacrr", pw::pacrr}, {"g6pbr", pw::g6pbr},
      {"precr1", pw::precr1}, {"precr2", pw::precr2}, {"roqimax", pw::roqimax}, {"pvts", pw::pvts},
      {"pacrs", pw::pacrs}, {"precs1", pw::precs1}, {"precs2", pw::precs2}, {"pidn0r", pw::pidn0r},
      {"pidn0s", pw::pidn0s}, {"xlv1", pw::xlv1}, {"pacrc", pw::pacrc}, {"pi", pw::pi}, {"pvtg", pw::pvtg},
      {"pacrg", pw::pacrg}, {"precg1", pw::precg1}, {"precg2", pw::precg2}, {"pidn0g", pw::pidn0g},
      {"rslopermax", pw::rslopermax}, {"rsloperbmax", pw::rsloperbmax}, {"rsloper2max", pw::rsloper2max},
      {"rsloper3max", pw::rsloper3max}, {"smdiv", pw::smdiv}, {"emdiv", pw::emdiv}, {"khdif", pw::khdif},
      {"kvdif", pw::kvdif}, {"epssm", pw::epssm}, {"dampcoef", pw::dampcoef}, {"zdamp", pw::zdamp},
  };
  if (!name || !value) return PYWRF_E_ARG;
  for (const Entry& e : table)
    if (strcmp(e.n, name) == 0) {
      *value = e.v;
      return 0;
    }
  return PYWRF_E_ARG;
}
