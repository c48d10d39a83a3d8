// Marching-cubes isosurface area of a cut cell (GeometryCalculator::get_area_MQ,
// geometry_calculator.cpp:284-318, with Polygonise_MQ::Polygonise and VertexInterp,
// polygonise.cpp:10-95), shared by the host and the GPU level-set builders so that both
// give the same bits.
//
// The cell's 3x3x3 corner samples C (Get_corner) split it into 2x2x2 sub-cubes, walked
// with the last index fastest.  A sub-cube's vertex v sits at sub-cube corner
// (kMcVert[v]): vertex 1 is the +y corner and vertex 3 the +x corner, as the reference
// numbers them.  Vertex v is inside when its value is below 0; the case index has bit v
// set for each inside vertex.  The triangle table is the published marching-cubes table
// (P. Bourke, "Polygonising a scalar field", the one the reference carries in
// polygonise.h), restated here as one 64-bit word per case: bits 0-3 hold the number of
// triangles, then each triangle's three edge indices, 4 bits each.  Edge e joins
// vertices kMcEdge[e]; a triangle vertex is the zero crossing on its edge, linearly
// interpolated (an end is taken as it is when its value is within 1e-15 of 0, or when
// the two values are within 1e-15 of each other).
//
// Triangle area: the reference evaluates 0.5 |p3 - p1| |p2 - p1| sin(acos(c)) with c the
// cosine of the angle at p1 (a NaN area counts 0).  Here sin(acos(c)) is formed as
// sqrt((1 - c)(1 + c)), the same quantity from correctly rounded operations only, so
// that host and device agree bit for bit; it differs from the library's sin(acos(c)) by
// a few ulp at most.  Sums run triangle by triangle in the reference's order.
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIP__)
#define SMG_MC_HD __host__ __device__
#else
#define SMG_MC_HD
#endif

// the 256 cases, four to a line
#define SMG_MC_TRI_TABLE \
  { \
    0x0000000000000000ull, 0x0000000000003801ull, 0x0000000000009101ull, 0x0000000001893812ull, \
    0x000000000000a211ull, 0x000000000a213802ull, 0x000000000920a292ull, 0x00000089a8a23823ull, \
    0x0000000000002b31ull, 0x0000000000b82b02ull, 0x000000000b320912ull, 0x000000b89b912b13ull, \
    0x0000000003ab1a32ull, 0x000000ab8a801a03ull, 0x0000009ab9b30933ull, 0x000000000b8aa892ull, \
    0x0000000000008741ull, 0x0000000004370342ull, 0x0000000007489102ull, 0x0000001371749143ull, \
    0x000000000748a212ull, 0x000000a214037433ull, 0x000000748209a293ull, 0x0004973727929a24ull, \
    0x0000000002b37482ull, 0x00000040242b74b3ull, 0x000000b327481093ull, 0x0001292b9b49b744ull, \
    0x000000487ab31a33ull, 0x0004b7401b41ab14ull, 0x00030bab9b098744ull, 0x000000ab99b4b743ull, \
    0x0000000000004591ull, 0x0000000003804592ull, 0x0000000000514502ull, 0x0000005135384583ull, \
    0x000000000459a212ull, 0x000000594a218033ull, 0x000000204245a253ull, 0x0008434535235a24ull, \
    0x000000000b324592ull, 0x000000594b802b03ull, 0x000000b325104503ull, 0x000584b828525124ull, \
    0x00000045931ab3a3ull, 0x000ab81a81805944ull, 0x00030bab5b050454ull, 0x000000b8aa858453ull, \
    0x0000000009758792ull, 0x0000003753590393ull, 0x0000007517108703ull, 0x0000000007533512ull, \
    0x00000021a7598793ull, 0x00037503505921a4ull, 0x00025a7585282084ull, 0x0000007533525a23ull, \
    0x0000002b39875973ull, 0x000b720292797594ull, 0x000751871810b324ull, 0x00000051771b12b3ull, \
    0x000b3a31a7588594ull, 0x0aba010b79050755ull, 0x07570805a30b0ab5ull, 0x0000000005b75ab2ull, \
    0x00000000000056a1ull, 0x0000000006a53802ull, 0x0000000006a51092ull, 0x0000006a58913813ull, \
    0x0000000001625612ull, 0x0000008036215613ull, 0x0000006206095693ull, 0x0008236252858954ull, \
    0x00000000056ab322ull, 0x00000056a02b80b3ull, 0x0000006a5b329103ull, 0x000b892b92916a54ull, \
    0x000000315356b363ull, 0x0006b51505b0b804ull, 0x0009505606306b34ull, 0x00000089bb969563ull, \
    0x0000000008746a52ull, 0x000000a563740343ull, 0x0000007486a50913ull, 0x00049737179156a4ull, \
    0x0000008741562163ull, 0x0007434036255214ull, 0x0006205605097484ull, 0x9626959234979375ull, \
    0x00000056a4872b33ull, 0x000b720242746a54ull, 0x0006a5b328749104ull, 0x6a54b7b492b91295ull, \
    0x0006b51535b37484ull, 0xb404b7b016b5b155ull, 0x74836b6305609505ull, 0x0009b7974b969564ull, \
    0x000000000a4694a2ull, 0x000000380a946a43ull, 0x00000004606a10a3ull, 0x000a164686181384ull, \
    0x0000004624219413ull, 0x0004629429218034ull, 0x0000000006244202ull, 0x0000006244282383ull, \
    0x00000032b46a94a3ull, 0x0006a4a94b822804ull, 0x000a164606102b34ull, 0x1b8b12184a161465ull, \
    0x00036b3196394694ull, 0x14641916b0181b85ull, 0x0000004600636b33ull, 0x00000000086b8462ull, \
    0x000000a98a876a73ull, 0x000a76a907a03704ull, 0x0000818717a176a4ull, 0x00000037117a76a3ull, \
    0x0007689818616214ull, 0x9373909761929625ull, 0x0000002066070873ull, 0x0000000002762372ull, \
    0x00076898a86ab324ull, 0x7a9a76790b727025ull, 0xb32a767a18710815ull, 0x00017616a71b12b4ull, \
    0x63136b6197686985ull, 0x00000000076b1902ull, 0x00006b0b36070874ull, 0x0000000000006b71ull, \
    0x000000000000b671ull, 0x00000000067b8032ull, 0x00000000067b9102ull, 0x00000067b1389183ull, \
    0x0000000007b621a2ull, 0x0000007b6803a213ull, 0x0000007b69a20923ull, 0x00089a38a3a27b64ull, \
    0x0000000007263272ull, 0x0000000260678073ull, 0x0000009107326723ull, 0x0006788916812614ull, \
    0x00000073171a67a3ull, 0x000801781a7167a4ull, 0x0007a69a0a707304ull, 0x0000009a88a7a673ull, \
    0x00000000068b4862ull, 0x000000640603b633ull, 0x000000109648b683ull, 0x00063b1393696494ull, \
    0x0000001a28b64863ull, 0x000640b60b03a214ull, 0x0009a2920b648b44ull, 0x36463b34923a39a5ull, \
    0x0000002642483283ull, 0x0000000002642402ull, 0x0008346424320914ull, 0x0000006422414913ull, \
    0x0001a66481683184ull, 0x00000040660a01a3ull, 0x39a9303a68343645ull, 0x0000000004a649a2ull, \
    0x000000000b675942ull, 0x00000067b5943803ull, 0x000000b670451053ull, 0x00051345343867b4ull, \
    0x000000b6721a4593ull, 0x000594380a217b64ull, 0x000204a24a45b674ull, 0x67b25a5234538435ull, \
    0x0000009452673273ull, 0x0007862606804594ull, 0x0000450516732634ull, 0x8515848127868265ull, \
    0x00073167161a4594ull, 0x459078701671a615ull, 0xa737a6a305a4a045ull, 0x000a84a458a7a674ull, \
    0x00000098b9b65963ull, 0x000590650360b634ull, 0x000b65510b508b04ull, 0x0000001355363b63ull, \
    0x00065b8b9b59a214ull, 0xa21965690b603b05ull, 0x52025a50865b58b5ull, 0x00035a3a25363b64ull, \
    0x0002832658259854ull, 0x0000002600696593ull, 0x8262838650818515ull, 0x0000000006126512ull, \
    0x698965683a616315ull, 0x00006505960a01a4ull, 0x000000000a658302ull, 0x00000000000065a1ull, \
    0x000000000b57a5b2ull, 0x00000003857ba5b3ull, 0x000000091ba57b53ull, 0x0001381897ba57a4ull, \
    0x00000015717b21b3ull, 0x000b275717213804ull, 0x0007b22097295794ull, 0x289823295b272575ull, \
    0x000000573532a523ull, 0x00052a5782580284ull, 0x0002a37353a51094ull, 0x25752a2781292895ull, \
    0x0000000005735312ull, 0x0000005711707803ull, 0x0000007355393093ull, 0x0000000007957892ull, \
    0x0000008ba8a54853ull, 0x00003bba50b54054ull, 0x00054aba8a489104ull, 0x41314943b54a4ba5ull, \
    0x0008548b25821524ull, 0xb151b2b543b0b405ull, 0x58b8545b29505205ull, 0x0000000003b25492ull, \
    0x000483543253a524ull, 0x0000000244252a53ull, 0x910854583a532a35ull, 0x0002492914252a54ull, \
    0x0000001533585483ull, 0x0000000005015402ull, 0x0005305093585484ull, 0x0000000000005491ull, \
    0x000000ba9b947b43ull, 0x000ba97b97943804ull, 0x000b470414b1ba14ull, 0x4bab474a18434135ull, \
    0x000219b294b97b44ull, 0x3801b2b197b94795ull, 0x00000004224b47b3ull, 0x00042343824b47b4ull, \
    0x000947732972a924ull, 0x70207872a4797a95ull, 0xa040a1a472a3a735ull, 0x0000000004782a12ull, \
    0x0000003177141943ull, 0x0001781807141944ull, 0x0000000003473042ull, 0x0000000000007841ull, \
    0x0000000008ba8a92ull, 0x000000a9bb939033ull, 0x000000ba88a0a103ull, 0x000000000a3ba132ull, \
    0x0000008b99b1b213ull, 0x0009b2921b939034ull, 0x000000000b08b202ull, 0x000000000000b231ull, \
    0x00000098aa828323ull, 0x0000000002902a92ull, 0x0008a1810a828324ull, 0x0000000000002a11ull, \
    0x0000000008198312ull, 0x0000000000001901ull, 0x0000000000008301ull, 0x0000000000000000ull \
  }

namespace smg {

// corners of vertex v of a sub-cube, as (x, y, z) offsets
SMG_MC_HD inline void mc_vert(int v, int& x, int& y, int& z) {
  // 0:(0,0,0) 1:(0,1,0) 2:(1,1,0) 3:(1,0,0), then the same four at z = 1
  const int b = v & 3;
  x = (b == 2 || b == 3) ? 1 : 0;
  y = (b == 1 || b == 2) ? 1 : 0;
  z = v >> 2;
}

// the two vertices of edge e: the four edges of the z = 0 face, the four of the z = 1
// face, then the four vertical edges
SMG_MC_HD inline void mc_edge(int e, int& a, int& b) {
  if (e < 8) {
    const int base = e & 4;
    a = base + (e & 3);
    b = base + ((e + 1) & 3);
  } else {
    a = e - 8;
    b = e - 4;
  }
}

// zero crossing of edge (p1, v1) - (p2, v2)
SMG_MC_HD inline void mc_interp(const double* p1, const double* p2, double v1, double v2, double* out) {
  const double* src = nullptr;
  if (std::fabs(v1) < 1.0e-15) src = p1;
  else if (std::fabs(v2) < 1.0e-15) src = p2;
  else if (std::fabs(v1 - v2) < 1.0e-15) src = p1;
  if (src) {
    out[0] = src[0];
    out[1] = src[1];
    out[2] = src[2];
    return;
  }
  const double mu = (0.0 - v1) / (v2 - v1);
  for (int c = 0; c < 3; c++) out[c] = p1[c] + mu * (p2[c] - p1[c]);
}

SMG_MC_HD inline double mc_tri_area(const double* p1, const double* p2, const double* p3) {
  const double u[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double w[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
  const double d1 = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const double d3 = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double c = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / (d3 * d1 + 1.e-20);
  const double a = 0.5 * (d3 * d1) * std::sqrt((1.0 - c) * (1.0 + c));
  return a != a ? 0.0 : a;
}

// The triangles of sub-cube (i, j, k) of a cell with corner samples C and cell size
// (dx, dy, dz): their areas in table order into out[0..4]; returns their number.
SMG_MC_HD inline int mc_subcube_areas(const double C[3][3][3], int i, int j, int k, double dx, double dy, double dz,
                                      const uint64_t* table, double* out) {
  double p[8][3], val[8];
  int idx = 0;
  for (int v = 0; v < 8; v++) {
    int x, y, z;
    mc_vert(v, x, y, z);
    const int vx = i + x, vy = j + y, vz = k + z;
    p[v][0] = vx * 0.5 * dx;
    p[v][1] = vy * 0.5 * dy;
    p[v][2] = vz * 0.5 * dz;
    val[v] = C[vx][vy][vz];
    if (val[v] < 0.0) idx |= 1 << v;
  }
  const uint64_t t = table[idx];
  const int n = int(t & 0xF);
  for (int q = 0; q < n; q++) {
    double tp[3][3];
    for (int s = 0; s < 3; s++) {
      const int e = int((t >> (4 + 4 * (3 * q + s))) & 0xF);
      int a, b;
      mc_edge(e, a, b);
      mc_interp(p[a], p[b], val[a], val[b], tp[s]);
    }
    out[q] = mc_tri_area(tp[0], tp[1], tp[2]);
  }
  return n;
}

// get_area_MQ: every sub-cube's triangles, added one by one in order
SMG_MC_HD inline double mc_cell_area(const double C[3][3][3], double dx, double dy, double dz, const uint64_t* table) {
  double area = 0.0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++) {
        double a[5];
        const int n = mc_subcube_areas(C, i, j, k, dx, dy, dz, table, a);
        for (int q = 0; q < n; q++) area += a[q];
      }
  return area;
}

}  // namespace smg
