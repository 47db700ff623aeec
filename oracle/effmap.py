"""TEST INFRASTRUCTURE ONLY — numpy restatement of the efficiency-map
finalisation (src/python/srf/preprocess/merge_map.py) and of the initial image
(src/python/srf/data/loader.py:28-37).  Images are numpy arrays of shape
(nx, ny, nz) in C order, the reference's L3 layout."""
import numpy as np


def make_meshes(grid, center, size):
    """RingBlock(size, grid, center, 0).get_meshes() (scanner/pet/block.py:77-108)
    with rad_z = 0 (identity rotation).  np.meshgrid's default 'xy' indexing
    orders points y-major, then x, then z."""
    grid = np.array(grid)
    size = np.array(size, dtype=np.float64)
    center = np.array(center, dtype=np.float64)
    interval = size / grid
    p_start = center - (size - interval) / 2
    p_end = center + (size - interval) / 2
    mr = [np.linspace(p_start[i], p_end[i], grid[i]) for i in range(3)]
    meshes = np.array(np.meshgrid(mr[0], mr[1], mr[2]))
    return np.transpose(np.reshape(meshes, (3, -1)))


def crop_image(inner_radius, axial_length, image, grid, center, size, crop_ratio):
    """merge_map.py:47-65."""
    half_height = axial_length / 2
    vox = make_meshes(grid, center, size)
    xy_dis = np.sqrt(vox[:, 0] ** 2 + vox[:, 1] ** 2)
    z_dis = np.abs(vox[:, 2])
    image = np.array(image).reshape((-1, 1)).copy()
    image[np.where(xy_dis > crop_ratio * inner_radius)] = 0
    image[np.where(z_dis >= half_height)] = 0
    return image.reshape((grid[0], grid[1], grid[2]))


def finalize_effmap(summed, inner_radius, axial_length, grid, center, size, crop_ratio=0.95):
    """merge_map.py:37-41: crop, divide by the max, invert where > 1e-6."""
    m = crop_image(inner_radius, axial_length, summed, grid, center, size, crop_ratio)
    m = m / np.max(m)
    m[m <= 1e-6] = 0
    m[m > 1e-6] = 1 / m[m > 1e-6]
    return m


def initial_image(inner_radius, axial_length, grid, center, size, crop_ratio=0.95):
    """MasterLoader.create_image, loader.py:28-37."""
    image = np.ones(grid, dtype=np.float32)
    return crop_image(inner_radius, axial_length, image, grid, center, size, crop_ratio)


def merge_effmap(maps, inner_radius, axial_length, grid, center, size, crop_ratio=0.95):
    """merge_map.py:8-44 without the files: maps[ir] is the (nx, ny, nz) map
    the reference saves as effmap_{ir}.npy and reads back transposed."""
    temp = np.asarray(maps[0]).T
    nb_image_layers = int(temp.shape[0])
    final_map = np.zeros(temp.shape)
    nb_rings = len(maps)
    for ir in range(nb_rings):
        temp = np.asarray(maps[ir]).T
        for jr in range(nb_rings - ir):
            final_map[jr:nb_image_layers, :, :] += temp[0:nb_image_layers - jr, :, :]
    final_map = crop_image(inner_radius, axial_length, final_map.T, grid, center, size,
                           crop_ratio)
    final_map = final_map / np.max(final_map)
    final_map[final_map <= 1e-6] = 0
    final_map[final_map > 1e-6] = 1 / final_map[final_map > 1e-6]
    return final_map


def merge_effmap_full(summed, inner_radius, axial_length, grid, center, size, crop_ratio=0.95):
    """merge_map.py:84-88 on the summed ordered-ring-pair maps."""
    final_map = crop_image(inner_radius, axial_length, np.asarray(summed, dtype=np.float64),
                           grid, center, size, crop_ratio)
    final_map = final_map / np.max(final_map)
    final_map[final_map > 1e-7] = 1 / final_map[final_map > 1e-7]
    return final_map
