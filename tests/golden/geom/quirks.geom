// Evaluation rules of the scene language, each written so that a different reading of the
// rule gives a different number.
chain_sub = 10 - 2 - 3
chain_div = 8/4/2
mixed = 2*3 + 4*5 - 6/3
power = 2^3^2
cmp_gt = 3 > 2
cmp_eq = (1 + 1) == 2
cmp_le = 2 <= 1
tern = (cmp_gt) ? 4 : 5
tern_plus = 1 + (0) ? 5 : 7
tern_else = (0) ? 1 : 2 + (1) ? 3 : 4
unknown_plus = undefined_name + 1
neg_name = -chain_sub
neg_num = -2.5 + 1
expo = 1e-3 + 2.5e+2
flags = true + false
lst = [1, 2, 3,4]
last = lst[-1]
first = lst[0]
sq = [x*x for x in range(4)]
sq_sum = sq[0] + sq[1] + sq[2] + sq[3]
lin = linspace(0, 1, 5)
lin_mid = lin[2]
fl = flatten([[1, 2], [3, [4, 5]]])
fl_n = fl[4]
trig = sin(0.5) + cos(0.5)*sqrt(2) + exp(0.1) + tan(0.2)
scaled = um_to_l(1.5)
fs = fs_to_t(10)
label = "a" + "b"
label_num = "v=" + 1.5
par = (((length)))
q1 = 1; q2 = q1 + 1
print("quirks: ", chain_sub, " ", tern_else)

Gaussian_source("Ez", 0.7, 2.0, 1.5, 0.25, start_time = 3.0, Box([0, 0, 0.5], [length, length, 0.5]))
CW_source("Ey", 0.9, 1.0, 0.5, 4.0, slowness = 2, Box([0, 0, 1], [length, length, 1]))
CW_source("Hx", 0.9, 1.0, 0.5, Box([0, 0, 1.5], [length, length, 1.5]))

monitors(locations = [vec(x, 1, 1) for x in linspace(0.5, 2.5, 3)])
monitors(locations = [[1, 1, 1], vec(um_to_l(0.25), par/2, 1.25)])

Composite(eps = 2.5, color = 7, susceptibilities = [[1.0 2.0 3.0, "lorentz"], [0.5, 0.1, 2.0, "drude"]], [
    Box([0, 0, 0], [1, 1, 1]),
    Sphere([2, 2, 2], 0.5),
])

Composite(eps = 1.5, combine_type = "intersect", [
    Union([
        Cylinder([1.5, 1.5, 1], 1.0, 0.5),
        Cylinder([1.5, 1.5, 2], 0.5, 0.25, 0.75)
    ]),
    Difference([
        Box([0, 0, 0], [3, 3, 3]),
        Plane([0, 0, 1], 2.0),
        Plane([0, 0, 0.5], [1, 0, 0.5], [0, 1, 0.75])
    ]),
    Rotate(0.3, [0, 0, 1], [Box([1, 1, 1], [2, 2, 2])]),
    Complement([Sphere([0.5, 0.5, 0.5], 0.25)]),
    Box([0.25, 0.25, 0.25], [2.75, 2.75, 2.75])
])
