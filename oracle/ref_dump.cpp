// Test-infrastructure harness, built only by oracle/Makefile from the reference's own
// headers where they lie (REF=/root/reference) into oracle/_ref/.  It is never shipped
// and never linked into the product.  It loads NAME.mpd with the reference's own reader
// and dumps, in the reference's own arithmetic and summation order (run it with
// OMP_NUM_THREADS=1), the step-0 quantities the restated oracle is pinned against:
//
//   PREFIX_keys.i32    cell key per particle      (CellOpt::build, cellOpt.h:530-556)
//   PREFIX_fpair.f64   pair forces, N x 3          (CellOpt::computeForce, cellOpt.h:712-926)
//                      (with -DFORCE_COUNT: per-particle neighbour counts, MD.h:838-845)
//   PREFIX_fchain.f64  CHAIN forces, N x 3         (Blob::doChainForce, system.h:1782-1866)
//   stdout             U_pair, U_chain, dU_pair(s), dU_chain(s), grid, as key=value lines
//
// usage: ref_dump NAME PREFIX [sx sy sz]
//
// The two configuration macros MD.cpp sets before its includes (MD.cpp:13,34):
#define CELL_SIZE_FAILURE
#define LOW_DENSITY
// CellOpt keeps its cell hash private; it is opened up for this harness only, after the
// standard headers are in (so the macro touches nothing but cellOpt.h itself).
#include "include/algorithms/functions.h"
#define private public
#include "include/algorithms/cellOpt.h"
#undef private
#include "include/MD.h"
#include "include/system.h"
#include <cstdio>
#include <cstdlib>
#include <vector>

static void dump(const char *prefix, const char *tag, const void *data, size_t bytes)
{
	std::string fn = std::string(prefix) + "_" + tag;
	FILE *f = fopen(fn.c_str(), "wb");
	if (!f) { fprintf(stderr, "cannot write %s\n", fn.c_str()); exit(2); }
	fwrite(data, 1, bytes, f);
	fclose(f);
}

int main(int argc, char **argv)
{
	if (argc != 3 && argc != 6) {
		fprintf(stderr, "usage: %s NAME PREFIX [sx sy sz]\n", argv[0]);
		return 1;
	}
	Blob<double> System;
	Script<double, Blob<double> > io(argv[1], std::ios::in, &System);
	io.read();
	io.close();

	const int N = System.readNParticles();
	threeVector<double> *a = System.getAccelerations();
	for (int i = 0; i < N; i++) a[i].x = a[i].y = a[i].z = 0;

	CellOpt<double, Potential<double>, Force<double> > pairs(System.getPositions(), a,
		System.getTwoBodyFconst(), System.getTwoBodyUconst(), N, System.readNTypes(),
		System.readSize(), System.readPeriodic(), System.readCutoff());
	pairs.build();

	std::vector<int> keys(N);
	for (int i = 0; i < N; i++) keys[i] = pairs.hashIndices[i].key;
	dump(argv[2], "keys.i32", keys.data(), sizeof(int) * N);

	pairs.computeForce();
	std::vector<double> buf(3 * (size_t)N);
	for (int i = 0; i < N; i++) { buf[3*i] = a[i].x; buf[3*i+1] = a[i].y; buf[3*i+2] = a[i].z; }
	dump(argv[2], "fpair.f64", buf.data(), sizeof(double) * buf.size());

	double Upair = pairs.computePotential();

	for (int i = 0; i < N; i++) a[i].x = a[i].y = a[i].z = 0;
	double Uchain = 0;
	for (int k = 0; k < System.readNMolecules(); k++) {
		if (System.getMolecule()[k].readType() == CHAIN) {
			System.doChainForce(k);
			Uchain += System.doChainPotential(k);
		}
	}
	for (int i = 0; i < N; i++) { buf[3*i] = a[i].x; buf[3*i+1] = a[i].y; buf[3*i+2] = a[i].z; }
	dump(argv[2], "fchain.f64", buf.data(), sizeof(double) * buf.size());

	printf("N=%d\n", N);
	printf("nCells=%d %d %d\n", pairs.nCells.x, pairs.nCells.y, pairs.nCells.z);
	printf("cellSize=%.17g %.17g %.17g\n", pairs.cellSize.x, pairs.cellSize.y, pairs.cellSize.z);
	printf("nFullCells=%d\n", pairs.nFullCells);
	printf("U_pair=%.17g\n", Upair);
	printf("U_chain=%.17g\n", Uchain);
	if (argc == 6) {
		threeVector<double> s;
		s.x = atof(argv[3]); s.y = atof(argv[4]); s.z = atof(argv[5]);
		double dUp = pairs.computeDPotential(s);
		double dUc = 0;
		for (int k = 0; k < System.readNMolecules(); k++)
			if (System.getMolecule()[k].readType() == CHAIN)
				dUc += System.doChainDPotential(k, s);
		printf("dU_pair=%.17g\n", dUp);
		printf("dU_chain=%.17g\n", dUc);
	}
	return 0;
}
