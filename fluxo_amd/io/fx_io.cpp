// fluxo_io: HOPR mesh files and state files in the reference's HDF5 formats
// (include/fluxo_io.h).  Host code over the HDF5 C library; nothing here touches the GPU.
#include "../../include/fluxo_io.h"

#include <hdf5.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// HDF5 handles closed on scope exit
struct H {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H() { if (id >= 0 && close) close(id); }
  H(const H &) = delete;
  H &operator=(const H &) = delete;
  operator hid_t() const { return id; }
  bool ok() const { return id >= 0; }
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

bool read_int_attr(hid_t loc, const char *name, int &v)
{
  if (H5Aexists(loc, name) <= 0) return false;
  H a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  return a.ok() && H5Aread(a, H5T_NATIVE_INT, &v) >= 0;
}

bool read_double_attr(hid_t loc, const char *name, double &v)
{
  if (H5Aexists(loc, name) <= 0) return false;
  H a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  return a.ok() && H5Aread(a, H5T_NATIVE_DOUBLE, &v) >= 0;
}

// first element of a fixed-length string attribute, trailing blanks removed
bool read_str_attr(hid_t loc, const char *name, std::string &s)
{
  if (H5Aexists(loc, name) <= 0) return false;
  H a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return false;
  H ft(H5Aget_type(a), H5Tclose);
  const size_t len = H5Tget_size(ft);
  H sp(H5Aget_space(a), H5Sclose);
  const hssize_t cnt = H5Sget_simple_extent_npoints(sp);
  if (len == 0 || cnt < 1 || H5Tis_variable_str(ft) > 0) return false;
  H mt(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(mt, len);
  std::vector<char> buf(len * (size_t)cnt + 1, '\0');
  if (H5Aread(a, mt, buf.data()) < 0) return false;
  s.assign(buf.data(), len);
  while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
  return true;
}

// dataset dimensions (C order) and contents
template <typename T>
bool read_dataset(hid_t file, const char *name, hid_t memtype, std::vector<hsize_t> &dims, std::vector<T> &data)
{
  if (H5Lexists(file, name, H5P_DEFAULT) <= 0) return false;
  H d(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return false;
  H sp(H5Dget_space(d), H5Sclose);
  const int rank = H5Sget_simple_extent_ndims(sp);
  if (rank < 1) return false;
  dims.assign(rank, 0);
  H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
  size_t n = 1;
  for (hsize_t x : dims) n *= (size_t)x;
  data.assign(n, T());
  return H5Dread(d, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.data()) >= 0;
}

bool write_str_attr(hid_t loc, const char *name, const std::vector<std::string> &v, size_t len)
{
  H t(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(t, len);
  H5Tset_strpad(t, H5T_STR_SPACEPAD);
  hsize_t n = v.size();
  H sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
  H a(H5Acreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return false;
  std::vector<char> buf(len * v.size(), ' ');
  for (size_t i = 0; i < v.size(); i++) std::memcpy(buf.data() + i * len, v[i].data(), std::min(len, v[i].size()));
  return H5Awrite(a, t, buf.data()) >= 0;
}

template <typename T>
bool write_num_attr(hid_t loc, const char *name, hid_t type, T v)
{
  hsize_t n = 1;
  H sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
  H a(H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return a.ok() && H5Awrite(a, type, &v) >= 0;
}

// HOPR ElemInfo / SideInfo columns (mesh_readin.f90:31-48) and BCType columns (:137-149)
enum { ELEM_TYPE = 0, ELEM_FIRSTSIDE = 2, ELEM_LASTSIDE = 3, ELEM_FIRSTNODE = 4, ELEM_LASTNODE = 5 };
enum { SIDE_TYPE = 0, SIDE_ID = 1, SIDE_NBELEM = 2, SIDE_FLIP = 3, SIDE_BCID = 4 };

struct Hopr {
  int nElems = 0, Ngeo = 0, gl = 0;
  std::vector<int> elemInfo, sideInfo, bcType;
  std::vector<double> nodes;
  size_t nBCs = 0;
};

int open_hopr(const char *file, Hopr &m, bool withData)
{
  quiet();
  if (!file) return fail(FXIO_EINVAL, "null file name");
  H f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(FXIO_EFILE, "cannot open HDF5 file '%s'", file);
  if (!read_int_attr(f, "Ngeo", m.Ngeo) || m.Ngeo < 1)
    return fail(FXIO_EFORMAT, "%s: attribute Ngeo missing or invalid", file);
  std::string nt = "EQUIDISTANT";
  if (H5Aexists(f, "OutputNodeType") > 0 && !read_str_attr(f, "OutputNodeType", nt))
    return fail(FXIO_EFORMAT, "%s: unreadable OutputNodeType", file);
  if (nt != "EQUIDISTANT" && nt != "GAUSS-LOBATTO")
    return fail(FXIO_EFORMAT, "%s: OutputNodeType must be EQUIDISTANT or GAUSS-LOBATTO, is '%s'", file, nt.c_str());
  m.gl = nt == "GAUSS-LOBATTO";
  std::vector<hsize_t> d;
  if (!read_dataset(f, "ElemInfo", H5T_NATIVE_INT, d, m.elemInfo) || d.size() != 2 || d[1] != 6)
    return fail(FXIO_EFORMAT, "%s: ElemInfo missing or not (nElems,6)", file);
  m.nElems = (int)d[0];
  if (!withData) return FXIO_OK;
  if (!read_dataset(f, "SideInfo", H5T_NATIVE_INT, d, m.sideInfo) || d.size() != 2 || d[1] != 5)
    return fail(FXIO_EFORMAT, "%s: SideInfo missing or not (nSides,5)", file);
  if (!read_dataset(f, "NodeCoords", H5T_NATIVE_DOUBLE, d, m.nodes) || d.size() != 2 || d[1] != 3)
    return fail(FXIO_EFORMAT, "%s: NodeCoords missing or not (nNodes,3)", file);
  if (H5Lexists(f, "BCType", H5P_DEFAULT) > 0) {
    if (!read_dataset(f, "BCType", H5T_NATIVE_INT, d, m.bcType) || d.size() != 2 || d[1] != 4)
      return fail(FXIO_EFORMAT, "%s: BCType not (nBCs,4)", file);
    m.nBCs = d[0];
  }
  return FXIO_OK;
}

// Connectivity of a HOPR mesh as the reference builds it: ReadMesh (mesh_readin.f90:299-440)
// reads each element's six sides, a big mortar side (negative SIDE_nbElemID: MortarType
// 1 = 1->4, 2 = 1->2 in eta, 3 = 1->2 in xi) followed by its 4 or 2 small virtual sides;
// setLocalSideIDs (prepare_mesh.f90:133-227) numbers the big mortar sides first and then
// the inner sides, in element / local side / mortar order; fillMeshInfo
// (prepare_mesh.f90:662-726) fills ElemToSide, SideToElem, MortarType and MortarInfo.
struct HSide {
  int elem = -1, loc = 0;      // owning element (0-based), local side 1..6 (0: virtual)
  int ind = 0, flip = 0;       // global side index, flip
  int nbElem = 0, bc = 0;      // SIDE_nbElemID, SIDE_BCID
  int mortarType = 0;          // >0 big mortar side, -10 small side of an element at a mortar
  int nMortars = 0, first = -1;  // small virtual sides: objects first .. first+nMortars-1
  int conn = -1, sideID = -1;
};

struct Conn {
  int E = 0, nSides = 0, nMortar = 0;
  std::vector<int> e2s, s2e, mortarType, mortarInfo;
};

int build_conn(const Hopr &m, Conn &c)
{
  const int E = m.nElems;
  const long long nSideRows = (long long)m.sideInfo.size() / 5;
  std::vector<HSide> sd;
  std::vector<int> loc(6 * E);  // object of each element's local side
  sd.reserve(6 * E);
  for (int e = 0; e < E; e++) {
    const int *ei = &m.elemInfo[6 * (size_t)e];
    const int type = ei[ELEM_TYPE];
    if (type != 108 && type != 118 && type != 208)
      return fail(FXIO_EUNSUPPORTED, "element %d has type %d; only hexahedra (108, 118, 208)", e + 1, type);
    long long row = ei[ELEM_FIRSTSIDE];
    auto next = [&](const int *&si) -> bool {
      if (row < 0 || row >= nSideRows || row >= ei[ELEM_LASTSIDE]) return false;
      si = &m.sideInfo[5 * (size_t)row++];
      return true;
    };
    for (int l = 0; l < 6; l++) {
      const int *si;
      if (!next(si)) return fail(FXIO_EFORMAT, "element %d: side rows end before local side %d", e + 1, l + 1);
      HSide a;
      a.elem = e; a.loc = l + 1; a.nbElem = si[SIDE_NBELEM]; a.bc = si[SIDE_BCID];
      if (si[SIDE_NBELEM] < 0) {
        a.mortarType = -si[SIDE_NBELEM];
        if (a.mortarType > 3) return fail(FXIO_EFORMAT, "element %d side %d: mortar type %d", e + 1, l + 1, a.mortarType);
        a.nMortars = a.mortarType == 1 ? 4 : 2;
      }
      if (si[SIDE_TYPE] < 0) a.mortarType = -10;
      if (a.mortarType <= 0) {
        a.ind = std::abs(si[SIDE_ID]);
        a.flip = si[SIDE_ID] > 0 ? 0 : si[SIDE_FLIP] % 10;
        if (a.flip < 0 || a.flip > 4) return fail(FXIO_EFORMAT, "element %d side %d: flip %d", e + 1, l + 1, a.flip);
      }
      loc[6 * e + l] = (int)sd.size();
      sd.push_back(a);
      if (a.nMortars > 0) {
        sd[loc[6 * e + l]].first = (int)sd.size();
        for (int im = 0; im < a.nMortars; im++) {
          if (!next(si)) return fail(FXIO_EFORMAT, "element %d side %d: small mortar side %d missing", e + 1, l + 1, im + 1);
          if (si[SIDE_ID] < 0)
            return fail(FXIO_EFORMAT, "element %d side %d: small mortar side %d is not oriented (flip 0)", e + 1, l + 1,
                        im + 1);
          HSide v;
          v.elem = e; v.ind = si[SIDE_ID]; v.nbElem = si[SIDE_NBELEM]; v.bc = si[SIDE_BCID];
          sd.push_back(v);
        }
      }
    }
    if (row != ei[ELEM_LASTSIDE])
      return fail(FXIO_EFORMAT, "element %d: %lld side rows, the six sides and their mortars use %lld", e + 1,
                  (long long)ei[ELEM_LASTSIDE] - ei[ELEM_FIRSTSIDE], row - ei[ELEM_FIRSTSIDE]);
  }
  // sides of element e, local side l, mortar im (0: the side itself), in the reference's order
  auto each = [&](auto &&f) {
    for (int e = 0; e < E; e++)
      for (int l = 0; l < 6; l++) {
        const int k = loc[6 * e + l];
        f(k);
        for (int im = 0; im < sd[k].nMortars; im++) f(sd[k].first + im);
      }
  };
  int err = FXIO_OK;
  // connections (mesh_readin.f90:368-440): the neighbour element's side or small virtual side
  // with the same index; periodic boundaries connect like inner sides
  each([&](int k) {
    HSide &a = sd[k];
    if (err) return;
    if (a.bc != 0) {
      if (a.bc < 1 || (size_t)a.bc > m.nBCs) { err = fail(FXIO_EFORMAT, "element %d: BC index %d", a.elem + 1, a.bc); return; }
      const int bcType = m.bcType[4 * (size_t)(a.bc - 1)];
      if (bcType != 1 && bcType != 100) {
        err = fail(FXIO_EUNSUPPORTED, "element %d is on boundary %d of type %d: only periodic boundaries are "
                   "supported", a.elem + 1, a.bc, bcType);
        return;
      }
    }
    if (a.mortarType > 0 || a.conn >= 0) return;
    if (a.nbElem < 1 || a.nbElem > E) {
      err = fail(FXIO_EUNSUPPORTED, "element %d side %d has no neighbour element: boundaries other than "
                 "periodic are not supported", a.elem + 1, a.loc);
      return;
    }
    const int nb = a.nbElem - 1;
    for (int l = 0; l < 6 && a.conn < 0; l++) {
      const int kb = loc[6 * nb + l];
      for (int im = -1; im < sd[kb].nMortars; im++) {
        const int kk = im < 0 ? kb : sd[kb].first + im;
        if (kk != k && sd[kk].mortarType <= 0 && sd[kk].ind == a.ind) { a.conn = kk; sd[kk].conn = k; break; }
      }
    }
    if (a.conn < 0) err = fail(FXIO_EFORMAT, "element %d side %d: neighbour side not found", a.elem + 1, a.loc);
  });
  if (err) return err;
  // side ids (setLocalSideIDs): big mortar sides 1..nMortar, then the inner sides
  int nMortar = 0;
  for (const HSide &a : sd) nMortar += a.mortarType > 0;
  int iMortar = 0, iInner = nMortar;
  each([&](int k) {
    HSide &a = sd[k];
    if (a.sideID >= 0) return;
    if (a.conn >= 0) a.sideID = sd[a.conn].sideID = ++iInner;
    else a.sideID = ++iMortar;  // every unconnected side is a big mortar side here
  });
  const int nS = iInner;
  c.E = E; c.nSides = nS; c.nMortar = nMortar;
  c.e2s.assign(12 * (size_t)E, 0);
  c.s2e.assign(5 * (size_t)nS, -1);
  c.mortarType.assign(2 * (size_t)nS, -1);
  c.mortarInfo.assign(10 * (size_t)nMortar, -1);
  // fillMeshInfo
  for (int e = 0; e < E; e++)
    for (int l = 0; l < 6; l++) {
      const HSide &a = sd[loc[6 * e + l]];
      c.e2s[12 * (size_t)e + 2 * l] = a.sideID;
      c.e2s[12 * (size_t)e + 2 * l + 1] = a.flip;
      int *t = &c.s2e[5 * (size_t)(a.sideID - 1)];
      if (a.flip == 0) {
        if (t[0] != -1) return fail(FXIO_EFORMAT, "side %d has two master elements", a.sideID);
        t[0] = e + 1; t[2] = l + 1;
      } else {
        if (t[1] != -1) return fail(FXIO_EFORMAT, "side %d has two slave elements", a.sideID);
        t[1] = e + 1; t[3] = l + 1; t[4] = a.flip;
      }
      if (a.nMortars > 0) {
        const int ms = a.sideID;  // 1..nMortar: its own index in MortarInfo
        c.mortarType[2 * (size_t)(ms - 1)] = a.mortarType;
        c.mortarType[2 * (size_t)(ms - 1) + 1] = ms;
        int *mi = &c.mortarInfo[10 * (size_t)(ms - 1)];
        mi[0] = ms; mi[1] = 0;
        for (int im = 0; im < a.nMortars; im++) {
          const HSide &v = sd[a.first + im];
          c.mortarType[2 * (size_t)(v.sideID - 1)] = 0;
          mi[2 * (im + 1)] = v.sideID;
          mi[2 * (im + 1) + 1] = v.flip;
        }
      }
    }
  // every inner side has its slave element; a conforming side its master, a small mortar side
  // none (its master data come from the big side)
  for (int s = nMortar; s < nS; s++) {
    const int *t = &c.s2e[5 * (size_t)s];
    if (t[1] < 1) return fail(FXIO_EFORMAT, "side %d lacks a slave element (orientation flags)", s + 1);
    if (t[0] < 1 && c.mortarType[2 * (size_t)s] != 0)
      return fail(FXIO_EFORMAT, "side %d lacks a master element (orientation flags)", s + 1);
  }
  return FXIO_OK;
}

}  // namespace

extern "C" {

const char *fxio_last_error(void) { return g_err; }

int fxio_mesh_info(const char *file, int *nElems, int *nSides, int *Ngeo, int *nodeTypeGL)
{
  Hopr m;
  Conn c;
  int rc = open_hopr(file, m, true);
  if (!rc) rc = build_conn(m, c);
  if (rc) return rc;
  if (nElems) *nElems = m.nElems;
  if (nSides) *nSides = c.nSides;
  if (Ngeo) *Ngeo = m.Ngeo;
  if (nodeTypeGL) *nodeTypeGL = m.gl;
  return FXIO_OK;
}

int fxio_mesh_read(const char *file, int *ElemToSide, int *SideToElem, double *NodeCoords)
{
  if (!ElemToSide || !SideToElem || !NodeCoords) return fail(FXIO_EINVAL, "null output array");
  Hopr m;
  Conn c;
  int rc = open_hopr(file, m, true);
  if (!rc) rc = build_conn(m, c);
  if (rc) return rc;
  const int E = m.nElems, ng = m.Ngeo + 1, ng3 = ng * ng * ng;
  const long long nNodeRows = (long long)m.nodes.size() / 3;
  for (int e = 0; e < E; e++) {
    const int *ei = &m.elemInfo[6 * (size_t)e];
    if (ei[ELEM_LASTNODE] - ei[ELEM_FIRSTNODE] != ng3 || ei[ELEM_FIRSTNODE] < 0 || ei[ELEM_LASTNODE] > nNodeRows)
      return fail(FXIO_EFORMAT, "element %d: node range [%d,%d) is not (Ngeo+1)^3 = %d nodes", e + 1,
                  ei[ELEM_FIRSTNODE], ei[ELEM_LASTNODE], ng3);
  }
  std::memcpy(ElemToSide, c.e2s.data(), sizeof(int) * c.e2s.size());
  std::memcpy(SideToElem, c.s2e.data(), sizeof(int) * c.s2e.size());
  // NodeCoords(3,0:Ngeo,0:Ngeo,0:Ngeo,nElems) (mesh_readin.f90:461-464)
  for (int e = 0; e < E; e++)
    std::memcpy(NodeCoords + (size_t)3 * ng3 * e, &m.nodes[3 * (size_t)m.elemInfo[6 * (size_t)e + ELEM_FIRSTNODE]],
                sizeof(double) * 3 * ng3);
  return FXIO_OK;
}

int fxio_mesh_mortars(const char *file, int *nMortarSides, int *MortarType, int *MortarInfo)
{
  if (!nMortarSides) return fail(FXIO_EINVAL, "null nMortarSides");
  Hopr m;
  Conn c;
  int rc = open_hopr(file, m, true);
  if (!rc) rc = build_conn(m, c);
  if (rc) return rc;
  *nMortarSides = c.nMortar;
  if (MortarType) std::memcpy(MortarType, c.mortarType.data(), sizeof(int) * c.mortarType.size());
  if (MortarInfo) std::memcpy(MortarInfo, c.mortarInfo.data(), sizeof(int) * c.mortarInfo.size());
  return FXIO_OK;
}

int fxio_state_write(const char *file, const char *projectName, const char *meshFile, int N, int nVar, int nElems,
                     double time, const char *const *varNames, const double *U)
{
  quiet();
  if (!file || !U || !varNames || N < 1 || nVar < 1 || nElems < 1) return fail(FXIO_EINVAL, "invalid argument");
  H f(H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(FXIO_EFILE, "cannot create '%s'", file);
  const std::string proj = projectName ? projectName : "", mesh = meshFile ? meshFile : "";
  // WriteHeader (hdf5_output.f90:685-702) and GenerateFileSkeleton (:559-614)
  bool ok = write_str_attr(f, "Program", {"Fluxo"}, 5) && write_str_attr(f, "File_Type", {"State"}, 5) &&
            write_str_attr(f, "Project_Name", {proj}, std::max<size_t>(1, proj.size())) &&
            write_num_attr(f, "File_Version", H5T_NATIVE_DOUBLE, 0.1);
  const hsize_t dims[5] = {(hsize_t)nElems, (hsize_t)N + 1, (hsize_t)N + 1, (hsize_t)N + 1, (hsize_t)nVar};
  {
    H sp(H5Screate_simple(5, dims, nullptr), H5Sclose);
    H d(H5Dcreate2(f, "DG_Solution", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    ok = ok && d.ok() && H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, U) >= 0;
  }
  std::vector<std::string> names;
  for (int v = 0; v < nVar; v++) names.push_back(varNames[v] ? varNames[v] : "");
  ok = ok && write_num_attr(f, "N", H5T_NATIVE_INT, N) && write_num_attr(f, "Time", H5T_NATIVE_DOUBLE, time) &&
       write_str_attr(f, "MeshFile", {mesh}, std::max<size_t>(1, mesh.size())) &&
       write_str_attr(f, "NodeType", {"GAUSS-LOBATTO"}, 255) && write_str_attr(f, "VarNames", names, 255) &&
       write_num_attr(f, "NComputation", H5T_NATIVE_INT, N);
  return ok ? FXIO_OK : fail(FXIO_EFILE, "writing '%s' failed", file);
}

int fxio_state_info(const char *file, int *N, int *nVar, int *nElems, double *time)
{
  quiet();
  if (!file) return fail(FXIO_EINVAL, "null file name");
  H f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(FXIO_EFILE, "cannot open HDF5 file '%s'", file);
  int n = 0;
  double t = 0.0;
  if (!read_int_attr(f, "N", n) || !read_double_attr(f, "Time", t))
    return fail(FXIO_EFORMAT, "%s: attributes N / Time missing", file);
  if (H5Lexists(f, "DG_Solution", H5P_DEFAULT) <= 0) return fail(FXIO_EFORMAT, "%s: no DG_Solution", file);
  H d(H5Dopen2(f, "DG_Solution", H5P_DEFAULT), H5Dclose);
  H sp(H5Dget_space(d), H5Sclose);
  hsize_t dims[5];
  if (H5Sget_simple_extent_ndims(sp) != 5) return fail(FXIO_EFORMAT, "%s: DG_Solution is not 5-dimensional", file);
  H5Sget_simple_extent_dims(sp, dims, nullptr);
  if ((int)dims[1] != n + 1 || dims[2] != dims[1] || dims[3] != dims[1])
    return fail(FXIO_EFORMAT, "%s: DG_Solution dimensions disagree with N=%d", file, n);
  if (N) *N = n;
  if (nVar) *nVar = (int)dims[4];
  if (nElems) *nElems = (int)dims[0];
  if (time) *time = t;
  return FXIO_OK;
}

int fxio_state_read(const char *file, int N, int nVar, int nElems, double *U)
{
  if (!U) return fail(FXIO_EINVAL, "null U");
  int n, nv, ne;
  double t;
  int rc = fxio_state_info(file, &n, &nv, &ne, &t);
  if (rc) return rc;
  H f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  std::string nt;
  if (!read_str_attr(f, "NodeType", nt)) return fail(FXIO_EFORMAT, "%s: NodeType missing", file);
  if (nt != "GAUSS-LOBATTO")
    return fail(FXIO_EUNSUPPORTED, "%s: node type %s (interpolation between node types is not supported)", file,
                nt.c_str());
  if (n != N) return fail(FXIO_EUNSUPPORTED, "%s: N=%d, computation N=%d (interpolation not supported)", file, n, N);
  if (nv != nVar || ne != nElems)
    return fail(FXIO_EINVAL, "%s: %d variables x %d elements, expected %d x %d", file, nv, ne, nVar, nElems);
  H d(H5Dopen2(f, "DG_Solution", H5P_DEFAULT), H5Dclose);
  if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, U) < 0)
    return fail(FXIO_EFILE, "%s: reading DG_Solution failed", file);
  return FXIO_OK;
}

}  // extern "C"
