// Which operand of corner + toX*r() + toY*r() takes the first draw when
// compiled with g++ as the reference is: vector3d.h declares the operators
// as non-member functions taking const references, so the expression is
// operator+(operator+(corner, operator*(toX, r())), operator*(toY, r())).
#include <cstdio>
struct vec { float x, y, z; };
struct pnt { float x, y, z; };
inline vec operator*(const vec& a, float f)
{
    vec r;
    r.x = f * a.x; r.y = f * a.y; r.z = f * a.z;
    return r;
}
inline pnt operator+(const pnt& a, const vec& b)
{
    pnt r;
    r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z;
    return r;
}
static int calls = 0;
static float draw() { return (float)(++calls); }
int main()
{
    pnt corner;
    corner.x = corner.y = corner.z = 0;
    vec toX, toY;
    toX.x = 1; toX.y = 0; toX.z = 0;
    toY.x = 0; toY.y = 1; toY.z = 0;
    const pnt p = corner + toX * draw() + toY * draw();
    std::printf("%s\n", (p.y == 1.0f && p.x == 2.0f) ? "toY first" : "toX first");
    return 0;
}
