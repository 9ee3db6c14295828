'use strict'
/*
 * The recorded hits on a detector as an illuminance map (optics/detmap.js detectorMap :104-213), from createOptics's
 * readHits(): the plane fitted to the hits (the normal is their least spread, by the covariance's smallest eigenvector;
 * within a degree of a world axis the map lies along the other two), square cells over the hits' box in the plane or a
 * given range, `cells` along the longer side, a cell's illuminance the flux landed in it over its area (lx for lm).
 *
 *   const m = detectorMap(hits, { cells: 64 })     hits: [{ pos, flux, dir, material }]
 *   m.E (Float64Array ny * nx, row 0 at the -v edge), m.max, m.meanLit, m.minLit, m.uniformity, m.total, m.cell (m)
 */
// the symmetric 3 x 3 matrix's eigenvectors by Jacobi rotations: [values ascending, vectors as columns]
function eigSym3(A) {
    const a = A.map(r => r.slice()), V = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    for (let sweep = 0; sweep < 50; sweep++) {
        let off = 0
        for (let p = 0; p < 3; p++) for (let q = p + 1; q < 3; q++) off += a[p][q] * a[p][q]
        if (off < 1e-30) break
        for (let p = 0; p < 3; p++) for (let q = p + 1; q < 3; q++) {
            if (Math.abs(a[p][q]) < 1e-300) continue
            const th = 0.5 * Math.atan2(2 * a[p][q], a[q][q] - a[p][p])
            const c = Math.cos(th), s = Math.sin(th)
            for (let k = 0; k < 3; k++) {            // columns p, q of a
                const akp = a[k][p], akq = a[k][q]
                a[k][p] = c * akp - s * akq; a[k][q] = s * akp + c * akq
            }
            for (let k = 0; k < 3; k++) {            // rows p, q
                const apk = a[p][k], aqk = a[q][k]
                a[p][k] = c * apk - s * aqk; a[q][k] = s * apk + c * aqk
            }
            for (let k = 0; k < 3; k++) {
                const vkp = V[k][p], vkq = V[k][q]
                V[k][p] = c * vkp - s * vkq; V[k][q] = s * vkp + c * vkq
            }
        }
    }
    const order = [0, 1, 2].sort((i, j) => a[i][i] - a[j][j])
    return [order.map(i => a[i][i]), order.map(i => [V[0][i], V[1][i], V[2][i]])]
}

function fitPlane(P) {
    const n = P.length, c = [0, 0, 0]
    for (const p of P) for (let k = 0; k < 3; k++) c[k] += p[k] / n
    const C = [[0, 0, 0], [0, 0, 0], [0, 0, 0]]
    for (const p of P) for (let i = 0; i < 3; i++) for (let j = 0; j < 3; j++) C[i][j] += (p[i] - c[i]) * (p[j] - c[j]) / Math.max(1, n - 1)
    const [, vec] = eigSym3(C)
    let normal = vec[0]
    const ax = [0, 1, 2].reduce((b, k) => Math.abs(normal[k]) > Math.abs(normal[b]) ? k : b, 0)
    if (Math.abs(normal[ax]) > Math.cos(Math.PI / 180)) {
        const e = k => [0, 1, 2].map(i => (i === k ? 1 : 0))
        const others = [0, 1, 2].filter(k => k !== ax)
        normal = e(ax).map(x => x * Math.sign(normal[ax]))
        return { centre: c, u: e(others[0]), v: e(others[1]), normal }
    }
    const u = vec[2]
    const v = [normal[1] * u[2] - normal[2] * u[1], normal[2] * u[0] - normal[0] * u[2], normal[0] * u[1] - normal[1] * u[0]]
    return { centre: c, u, v, normal }
}

function detectorMap(hits, { cells = 64, range = null, metresPerUnit = 1 } = {}) {
    if (!hits.length) return null
    const { centre, u, v, normal } = fitPlane(hits.map(h => h.pos))
    const dot = (p, w) => (p[0] - centre[0]) * w[0] + (p[1] - centre[1]) * w[1] + (p[2] - centre[2]) * w[2]
    const A = hits.map(h => dot(h.pos, u)), B = hits.map(h => dot(h.pos, v))
    const [a0, a1] = range ? range.a : [Math.min(...A), Math.max(...A)]
    const [b0, b1] = range ? range.b : [Math.min(...B), Math.max(...B)]
    const side = Math.max(a1 - a0, b1 - b0) / cells
    const nx = Math.max(1, Math.ceil((a1 - a0) / side - 1e-4)), ny = Math.max(1, Math.ceil((b1 - b0) / side - 1e-4))
    const flux = new Float64Array(nx * ny), count = new Uint32Array(nx * ny)
    let total = 0
    hits.forEach((h, k) => {
        if (A[k] < a0 || A[k] > a1 || B[k] < b0 || B[k] > b1) return
        const ix = Math.min(nx - 1, Math.max(0, Math.floor((A[k] - a0) / side)))
        const iy = Math.min(ny - 1, Math.max(0, Math.floor((B[k] - b0) / side)))
        flux[iy * nx + ix] += h.flux; count[iy * nx + ix]++; total += h.flux
    })
    const area = (side * metresPerUnit) ** 2
    const E = flux.map(f => f / area)
    const lit = []
    for (let i = 0; i < E.length; i++) if (count[i]) lit.push(E[i])
    const mean = lit.reduce((s, x) => s + x, 0) / lit.length
    return { E, count, nx, ny, cell: side * metresPerUnit, centre, u, v, normal, range: { a: [a0, a1], b: [b0, b1] }, total,
             max: Math.max(...lit), meanLit: mean, minLit: Math.min(...lit), uniformity: mean > 0 ? Math.min(...lit) / mean : null }
}

module.exports = { detectorMap, fitPlane }
