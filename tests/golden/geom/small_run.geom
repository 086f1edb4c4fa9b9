// A small two-electrode gap with a Drude metal, a sheet source and a line of monitors across
// the gap, for end-to-end driver runs.
gap = um_to_l(0.1)
mid = length/2
half = 0.25
pwidth = 1.5
Gaussian_source("Ex", 0.8, 1.0, pwidth, 0.3, cutoff=4, Box([0,0,0.75], [length,length,0.75]))
monitors(locations = [vec(x, mid, mid) for x in linspace(mid - gap, mid + gap, 3)])
Composite(eps = 2.0, susceptibilities = [[1e-10, 0.05, 4.0e21, "drude"]], [
    Box([0, 0, mid - half], [mid - gap/2, length, mid + half]),
    Box([mid + gap/2, 0, mid - half], [length, length, mid + half]),
])
