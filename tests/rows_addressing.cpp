// Host check of the fused MD step's rows-path addressing (rows_at_lane,
// rows_site, blk_off in su3_dev.h): for every site of lattices that take the
// path (L3 | 64, (L2 L3) % 64 == 0), every link direction and every shift by
// one or two of +/- e0, e1, e3, the block and slot found from the site's block
// index must give the element offset site_off gives for the shifted site,
// and the block's x0/x1 coordinates and index must be the same for all 64 of
// its sites (the kernel takes them from the first lane).  Exit status 0 when
// every offset agrees.
#include "../mesh_amd/csrc/su3_dev.h"
#include <cstdio>

using namespace mesh_dev;

static Geo make_geo(int l0, int l1, int l2, int l3)
{
	Geo g{};
	g.nd = 4;
	g.L[0] = l0;
	g.L[1] = l1;
	g.L[2] = l2;
	g.L[3] = l3;
	g.st[3] = 1;
	g.st[2] = l3;
	g.st[1] = l2 * l3;
	g.st[0] = l1 * l2 * l3;
	g.V = (long)l0 * l1 * l2 * l3;
	g.stride = (g.V + 255) / 256 * 256;
	return g;
}

static void coords_of(Geo const &g, long x, int *c)
{
	for (int m = 3; m >= 0; --m)
	{
		c[m] = (int)(x % g.L[m]);
		x /= g.L[m];
	}
}

int main()
{
	const int geos[][4] = {{16, 16, 16, 16}, {8, 8, 8, 8}, {4, 8, 4, 16}, {16, 4, 8, 8},
	                       {6, 10, 2, 32},   {4, 4, 16, 4}, {12, 8, 16, 64}, {3, 5, 4, 16}};
	long bad = 0, n = 0;
	for (auto &G : geos)
	{
		Geo g = make_geo(G[0], G[1], G[2], G[3]);
		if (64 % g.L[3] != 0 || (g.L[2] * g.L[3]) % 64 != 0)
		{
			std::printf("geometry %d %d %d %d does not take the rows path\n", G[0], G[1], G[2], G[3]);
			return 2;
		}
		for (long x = 0; x < g.V; ++x)
		{
			int c[4], cb[4];
			coords_of(g, x, c);
			coords_of(g, x & ~63L, cb);
			RowsAt a = rows_at_lane(g, x, c);
			if (cb[0] != c[0] || cb[1] != c[1] || a.lb != (int)((x & ~63L) >> 6))
				++bad;
			for (int a1 : {0, 1, 3})
				for (int s1 : {-1, 1})
					for (int a2 : {-1, 0, 1, 3})
						for (int s2 : {-1, 1})
						{
							if (a2 == a1 || (a2 < 0 && s2 > 0))
								continue;
							long xs = shift_c(g, x, a1, c[a1], s1);
							if (a2 >= 0)
								xs = shift_c(g, xs, a2, c[a2], s2);
							BlkSlot b = rows_site(g, a, a1, s1, a2, s2);
							for (int f = 0; f < 4; ++f, ++n)
								if (blk_off(g, f, b.blk, b.l) != site_off(g, f, xs))
								{
									if (bad < 5)
										std::printf("L %d %d %d %d x %ld shift (%d,%d)(%d,%d) f %d\n", G[0], G[1], G[2],
										            G[3], x, a1, s1, a2, s2, f);
									++bad;
								}
						}
		}
	}
	std::printf("%ld offsets, %ld wrong\n", n, bad);
	return bad != 0;
}
